// Host code of the library and of the C restatement, built with AddressSanitizer
// and UndefinedBehaviorSanitizer (tests/test_sanitizers.py): the template-bank
// generator (own Fresnel integrals, fdas_templates.cpp), the SIGPROC reader with
// its error paths (fdas_io.cpp), and the C restatement's FOP and harmonic
// summing (oracle/fdas_cpu.c) on small ragged inputs with every tile size.
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "fdas.h"

extern "C" {
int fdas_cpu_fop(const float* x_c64, int n, const float* taps_c64, int t_count, int taps, int tile, float* fop,
                 int n_threads);
int fdas_cpu_hsum(const float* fop, int t_count, int n, int n_harm, const float* thr, int cap, uint32_t* n_detected,
                  uint32_t* key_out, float* pwr_out, uint32_t* n_out, int n_threads);
}

static int failures = 0;
#define CHECK(c)                                                 \
  do {                                                           \
    if (!(c)) {                                                  \
      std::printf("FAIL %s:%d: %s\n", __FILE__, __LINE__, #c);   \
      ++failures;                                                \
    }                                                            \
  } while (0)

static void put_key(std::string& h, const char* k) {
  const int32_t n = int32_t(std::strlen(k));
  h.append(reinterpret_cast<const char*>(&n), 4);
  h.append(k);
}

static std::string tim_file(const char* path, int32_t data_type, int32_t nbits, bool end, const std::vector<float>& x) {
  std::string h;
  put_key(h, "HEADER_START");
  put_key(h, "data_type");
  h.append(reinterpret_cast<const char*>(&data_type), 4);
  put_key(h, "nbits");
  h.append(reinterpret_cast<const char*>(&nbits), 4);
  put_key(h, "tsamp");
  const double ts = 6.4e-5;
  h.append(reinterpret_cast<const char*>(&ts), 8);
  if (end) put_key(h, "HEADER_END");
  FILE* f = std::fopen(path, "wb");
  std::fwrite(h.data(), 1, h.size(), f);
  std::fwrite(x.data(), sizeof(float), x.size(), f);
  std::fclose(f);
  return path;
}

int main(int argc, char** argv) {
  const std::string dir = argc > 1 ? argv[1] : ".";
  // template bank: the reference's drifts, 85 x 421, and a few odd shapes
  std::vector<double> drifts(85);
  CHECK(fdas_reference_drifts(85, 350.0, 500.0, 600.0, 299792458.0, drifts.data()) == FDAS_OK);
  std::vector<float> bank(85 * 421 * 2);
  CHECK(fdas_make_drift_bank(drifts.data(), 85, 421, bank.data()) == FDAS_OK);
  double energy = 0.0;
  for (int m = 0; m < 421; ++m) energy += double(bank[2 * (10 * 421 + m)]) * bank[2 * (10 * 421 + m)] +
                                          double(bank[2 * (10 * 421 + m) + 1]) * bank[2 * (10 * 421 + m) + 1];
  CHECK(energy > 0.99 && energy < 1.01);
  const double odd[3] = {-0.3, 0.0, 7.25};
  std::vector<float> small(3 * 20 * 2);
  CHECK(fdas_make_drift_bank(odd, 3, 20, small.data()) == FDAS_OK);
  CHECK(fdas_make_drift_bank(odd, 3, 2, small.data()) != FDAS_OK);  // 7.25 needs more than 2 taps
  CHECK(fdas_make_drift_bank(nullptr, 3, 20, small.data()) != FDAS_OK);

  // SIGPROC reader: good file, truncated header, wrong format, missing file, short buffer
  std::vector<float> x(1000);
  for (size_t i = 0; i < x.size(); ++i) x[i] = float(std::sin(0.01 * double(i)));
  const std::string good = tim_file((dir + "/good.tim").c_str(), 2, 32, true, x);
  uint64_t n = 0;
  double ts = 0.0;
  CHECK(fdas_read_tim(good.c_str(), 1, nullptr, 0, &n, &ts) == FDAS_OK && n == x.size());
  std::vector<float> got(n);
  CHECK(fdas_read_tim(good.c_str(), 1, got.data(), got.size(), &n, &ts) == FDAS_OK && got == x && ts == 6.4e-5);
  std::vector<float> tiny(10);
  fdas_read_tim(good.c_str(), 1, tiny.data(), tiny.size(), &n, &ts);  // either an error or a bounded read
  CHECK(fdas_read_tim(tim_file((dir + "/noend.tim").c_str(), 2, 32, false, x).c_str(), 1, got.data(), got.size(), &n,
                      &ts) != FDAS_OK);
  CHECK(fdas_read_tim(tim_file((dir + "/bits.tim").c_str(), 2, 8, true, x).c_str(), 1, got.data(), got.size(), &n,
                      &ts) != FDAS_OK);
  CHECK(fdas_read_tim((dir + "/missing.tim").c_str(), 1, got.data(), got.size(), &n, &ts) != FDAS_OK);
  CHECK(std::strlen(fdas_io_last_error()) > 0);

  // C restatement: FOP then harmonic summing, ragged N, all tile sizes, a cap that binds
  const int T = 9, taps = 169;
  std::vector<double> d9(T);
  for (int t = 0; t < T; ++t) d9[t] = 2.0 * (t - T / 2) * 8.0;
  std::vector<float> b9(T * taps * 2);
  CHECK(fdas_make_drift_bank(d9.data(), T, taps, b9.data()) == FDAS_OK);
  for (int nb : {1, 419, 1629, 5003}) {
    std::vector<float> spec(2 * nb);
    for (int i = 0; i < 2 * nb; ++i) spec[i] = float(std::cos(0.37 * i) + 0.1 * std::sin(1.3 * i * i));
    for (int tile : {1024, 2048, 4096}) {
      std::vector<float> fop(size_t(T) * nb);
      CHECK(fdas_cpu_fop(spec.data(), nb, b9.data(), T, taps, tile, fop.data(), 2) == 0);
      const int H = 16, cap = 7;
      std::vector<float> thr(H);
      for (int k = 0; k < H; ++k) thr[k] = 0.5f * float(k + 1);
      std::vector<uint32_t> det(H), keys(H * cap), nout(H);
      std::vector<float> pw(H * cap);
      CHECK(fdas_cpu_hsum(fop.data(), T, nb, H, thr.data(), cap, det.data(), keys.data(), pw.data(), nout.data(), 2) ==
            0);
    }
  }
  CHECK(fdas_cpu_fop(nullptr, 10, b9.data(), T, taps, 1024, nullptr, 1) != 0);
  if (failures) {
    std::printf("%d failures\n", failures);
    return 1;
  }
  std::printf("host sanitize test ok\n");
  return 0;
}
