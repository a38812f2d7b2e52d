// Host-side driver of the reference's test flows against the drop-in C++ class.
//
// Follows fdas/src/host/main_gtest.cpp: value-parameterised over test-vector
// directories, each holding input.npy, thresholds.npy, fop_ref.npy,
// det_loc_ref.npy and det_pwr_ref.npy (written by the oracle), plus the
// template file given with -t (the reference's fdas_templates_85_350_ft_p4.npy
// in FFT order).  Runs FT_Convolution, Harmonic_Summing, FDAS_steps (the
// single-step tests, main_gtest.cpp:235-311), drive_serially and
// drive_pipelined in the three buffer modes (:315-465), with the candidate
// validator of :192-228.  Prints one line per test and exits non-zero on any
// failure.
//
// Tolerances: the FOP is compared relative to its maximum (1e-5) because the
// references are FP64; a device candidate missing from the reference list is
// accepted only when its power lies within 1e-4 (relative) of the plane's
// threshold, the FP32-vs-FP64 band.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "FDAS.hpp"

using fdas_amd::FDAS;
using fdas_amd::float2_t;

namespace {

// Minimal .npy reader: little-endian, C order, dtypes <f4, <c8, <u4; `per` values of T per element.
template <typename T>
bool load_npy(const std::string& path, const char* want, std::vector<T>& out, std::vector<size_t>& shape,
              size_t per = 1) {
  std::ifstream f(path, std::ios::binary);
  if (!f) return false;
  char magic[6];
  f.read(magic, 6);
  if (std::memcmp(magic, "\x93NUMPY", 6) != 0) return false;
  unsigned char ver[2];
  f.read(reinterpret_cast<char*>(ver), 2);
  uint32_t hlen = 0;
  if (ver[0] == 1) {
    uint16_t h;
    f.read(reinterpret_cast<char*>(&h), 2);
    hlen = h;
  } else {
    f.read(reinterpret_cast<char*>(&hlen), 4);
  }
  std::string header(hlen, ' ');
  f.read(&header[0], hlen);
  if (header.find(want) == std::string::npos || header.find("'fortran_order': False") == std::string::npos)
    return false;
  shape.clear();
  const size_t a = header.find('('), b = header.find(')');
  std::stringstream ss(header.substr(a + 1, b - a - 1));
  std::string tok;
  size_t count = 1;
  while (std::getline(ss, tok, ',')) {
    if (tok.find_first_not_of(" ") == std::string::npos) continue;
    shape.push_back(std::stoul(tok));
    count *= shape.back();
  }
  out.resize(count * per);
  f.read(reinterpret_cast<char*>(out.data()), std::streamsize(count * per * sizeof(T)));
  return bool(f);
}

struct Vectors {
  std::vector<float2_t> input, templates;
  std::vector<float> thresholds, fop_ref, pwr_ref;
  std::vector<uint32_t> loc_ref;
  std::vector<size_t> fop_shape;
};

int g_failures = 0;
int g_runs = 4;

void report(const std::string& name, bool ok, const std::string& why = "") {
  std::printf("[%s] %s%s%s\n", ok ? "  OK  " : " FAIL ", name.c_str(), why.empty() ? "" : ": ", why.c_str());
  std::fflush(stdout);
  if (!ok) ++g_failures;
}

#define REQUIRE(cond)                                     \
  do {                                                    \
    if (!(cond)) {                                        \
      report(name, false, #cond " [line " + std::to_string(__LINE__) + "]"); \
      return;                                             \
    }                                                     \
  } while (0)

// main_gtest.cpp:192-228, with the FP32-vs-FP64 threshold band.
std::string validate_candidates(const Vectors& v, std::vector<uint32_t> loc, std::vector<float> pwr, int c) {
  const uint32_t invalid = uint32_t((2 * c + 1) & 0x7f) << 22;  // encode(1, c+1, 0), gen_info.h:69
  size_t kept = 0;
  for (size_t i = 0; i < loc.size(); ++i)
    if (loc[i] != invalid) {
      loc[kept] = loc[i];
      pwr[kept] = pwr[i];
      ++kept;
    }
  loc.resize(kept);
  pwr.resize(kept);
  if (kept < 1) return "no candidates";
  if (kept > v.loc_ref.size()) return "more candidates than the reference list";
  for (size_t i = 0; i < kept; ++i) {
    auto it = std::find(v.loc_ref.begin(), v.loc_ref.end(), loc[i]);
    const uint32_t h = ((loc[i] >> 29) & 7u) + 1;
    const float thr = v.thresholds[h - 1];
    if (it == v.loc_ref.end()) {
      if (std::fabs(pwr[i] - thr) > 1e-4f * thr) {
        std::ostringstream os;
        os << "INVALID harmonic=" << h << " template=" << int((loc[i] >> 22) & 0x7f) - c
           << " freq.bin=" << (loc[i] & 0x3fffff) << " power=" << pwr[i];
        return os.str();
      }
      continue;
    }
    const float ref = v.pwr_ref[size_t(it - v.loc_ref.begin())];
    if (std::fabs(pwr[i] - ref) > 1e-4f * std::fabs(ref)) return "power mismatch";
  }
  return "";
}

void validate_fop(const std::string& name, const Vectors& v, const std::vector<float>& fop, const std::vector<int>& rows,
                  size_t n) {
  double mx = 0, err = 0;
  for (int r : rows)
    for (size_t f = 0; f < n; ++f) {
      mx = std::max(mx, double(v.fop_ref[r * n + f]));
      err = std::max(err, std::fabs(double(fop[r * n + f]) - double(v.fop_ref[r * n + f])));
    }
  report(name, err <= 1e-5 * mx, "max |dFOP| / max FOP = " + std::to_string(err / mx));
}

std::vector<int> rows_of(FDAS::FOPPart p, int c) {
  std::vector<int> r;
  const int lo = p == FDAS::PositiveAccelerations ? c : 0;
  const int hi = p == FDAS::NegativeAccelerations ? c : 2 * c;
  for (int i = lo; i <= hi; ++i) r.push_back(i);
  return r;
}

void test_ft_convolution(const Vectors& v) {
  const std::string name = "FT_Convolution";
  FDAS pipeline(std::cerr);
  fdas_config cfg = pipeline.config();
  cfg.flags |= FDAS_FLAG_REFERENCE_FFT_ORDER;
  pipeline.configure(cfg);
  REQUIRE(pipeline.initialise_accelerator("", FDAS::choose_first_platform, FDAS::choose_accelerator_devices,
                                          uint32_t(v.input.size())));
  REQUIRE(pipeline.get_templates_sz() == v.templates.size());
  REQUIRE(pipeline.upload_templates(v.templates.data(), FDAS::A));
  REQUIRE(pipeline.perform_input_tiling(v.input.data(), FDAS::A));
  REQUIRE(pipeline.perform_ft_convolution(FDAS::AllAccelerations, FDAS::A));
  std::vector<float2_t> tiles(pipeline.get_tiles_sz());
  REQUIRE(pipeline.retrieve_tiles(tiles.data(), FDAS::A));
  std::vector<float> fop(pipeline.get_fop_sz());
  REQUIRE(pipeline.retrieve_FOP(fop.data(), FDAS::A));
  validate_fop(name, v, fop, rows_of(FDAS::AllAccelerations, 42), v.input.size());
}

void test_harmonic_summing(const Vectors& v) {
  const std::string name = "Harmonic_Summing";
  FDAS pipeline(std::cerr);
  REQUIRE(pipeline.initialise_accelerator("", FDAS::choose_first_platform, FDAS::choose_accelerator_devices,
                                          uint32_t(v.input.size())));
  REQUIRE(v.fop_ref.size() == pipeline.get_fop_sz());
  REQUIRE(pipeline.inject_FOP(v.fop_ref.data(), FDAS::A));
  REQUIRE(pipeline.perform_harmonic_summing(v.thresholds.data(), FDAS::NegativeAccelerations, FDAS::A));
  std::vector<uint32_t> loc(pipeline.get_candidate_list_sz());
  std::vector<float> pwr(pipeline.get_candidate_list_sz());
  REQUIRE(pipeline.retrieve_candidates(loc.data(), pwr.data(), FDAS::A));
  // the reference list covers all rows; negative-part candidates are a subset of it
  const std::string why = validate_candidates(v, loc, pwr, 42);
  report(name, why.empty(), why);
}

void test_steps(const Vectors& v) {
  const std::string name = "FDAS_steps";
  FDAS pipeline(std::cerr);
  fdas_config cfg = pipeline.config();
  cfg.flags |= FDAS_FLAG_REFERENCE_FFT_ORDER;
  pipeline.configure(cfg);
  REQUIRE(pipeline.initialise_accelerator("", FDAS::choose_first_platform, FDAS::choose_accelerator_devices,
                                          uint32_t(v.input.size())));
  REQUIRE(pipeline.get_templates_sz() == v.templates.size());
  REQUIRE(pipeline.get_thresholds_sz() == v.thresholds.size());
  REQUIRE(pipeline.upload_templates(v.templates.data(), FDAS::B));
  REQUIRE(pipeline.upload_thresholds(v.thresholds.data(), FDAS::B));
  REQUIRE(pipeline.perform_input_tiling(v.input.data(), FDAS::B));
  REQUIRE(pipeline.perform_ft_convolution(FDAS::PositiveAccelerations, FDAS::B));
  REQUIRE(pipeline.perform_harmonic_summing(v.thresholds.data(), FDAS::PositiveAccelerations, FDAS::B));
  std::vector<uint32_t> loc(pipeline.get_candidate_list_sz());
  std::vector<float> pwr(pipeline.get_candidate_list_sz());
  REQUIRE(pipeline.retrieve_candidates(loc.data(), pwr.data(), FDAS::B));
  const std::string why = validate_candidates(v, loc, pwr, 42);
  report(name, why.empty(), why);
}

void drive_serially(const Vectors& v, FDAS::BufferMode mode, const std::string& name) {
  FDAS pipeline(std::cerr);
  fdas_config cfg = pipeline.config();
  cfg.flags |= FDAS_FLAG_REFERENCE_FFT_ORDER;
  pipeline.configure(cfg);
  REQUIRE(pipeline.initialise_accelerator("", FDAS::choose_first_platform, FDAS::choose_accelerator_devices,
                                          uint32_t(v.input.size()), mode));
  REQUIRE(pipeline.get_templates_sz() == v.templates.size());
  REQUIRE(pipeline.upload_templates(v.templates.data(), FDAS::A));
  std::vector<uint32_t> loc(pipeline.get_candidate_list_sz());
  std::vector<float> pwr(pipeline.get_candidate_list_sz());
  for (int i = 0; i < g_runs; ++i) {
    const FDAS::FOPPart which = (i & 1) == 0 ? FDAS::PositiveAccelerations : FDAS::NegativeAccelerations;
    REQUIRE(pipeline.launch(v.input.data(), v.thresholds.data(), loc.data(), pwr.data(), which, FDAS::A));
    REQUIRE(pipeline.wait(FDAS::A));
    pipeline.print_stats(FDAS::A, i == 0);
    pipeline.print_events(FDAS::A);
  }
  const std::string why = validate_candidates(v, loc, pwr, 42);
  report(name, why.empty(), why);
}

void drive_pipelined(const Vectors& v, FDAS::BufferMode mode, const std::string& name) {
  FDAS pipeline(std::cerr);
  fdas_config cfg = pipeline.config();
  cfg.flags |= FDAS_FLAG_REFERENCE_FFT_ORDER;
  pipeline.configure(cfg);
  REQUIRE(pipeline.initialise_accelerator("", FDAS::choose_first_platform, FDAS::choose_accelerator_devices,
                                          uint32_t(v.input.size()), mode, true));
  REQUIRE(pipeline.get_templates_sz() == v.templates.size());
  REQUIRE(pipeline.upload_templates(v.templates.data(), FDAS::A));
  REQUIRE(pipeline.upload_templates(v.templates.data(), FDAS::B));
  std::vector<uint32_t> loc[2];
  std::vector<float> pwr[2];
  for (int ab = 0; ab < 2; ++ab) {
    loc[ab].resize(pipeline.get_candidate_list_sz());
    pwr[ab].resize(pipeline.get_candidate_list_sz());
  }
  const FDAS::FOPPart which[2] = {FDAS::PositiveAccelerations, FDAS::NegativeAccelerations};
  for (int i = 0; i < g_runs + 1; ++i) {
    const FDAS::BufferSet ab = (i & 1) == 0 ? FDAS::A : FDAS::B;
    const FDAS::BufferSet prev = (i & 1) == 0 ? FDAS::B : FDAS::A;
    if (i < g_runs)
      REQUIRE(pipeline.launch(v.input.data(), v.thresholds.data(), loc[ab].data(), pwr[ab].data(), which[ab], ab));
    else
      REQUIRE(pipeline.end_pipeline(ab));
    if (i > 0) {
      REQUIRE(pipeline.wait(prev));
      pipeline.print_stats(prev, i - 1 == 0);
      pipeline.print_events(prev);
    }
  }
  for (int ab = 0; ab < 2; ++ab) {
    const std::string why = validate_candidates(v, loc[ab], pwr[ab], 42);
    report(name + (ab == 0 ? "[A]" : "[B]"), why.empty(), why);
  }
}

bool load_vectors(const std::string& dir, const std::string& templates_file, Vectors& v) {
  std::vector<size_t> shape;
  std::vector<float> c;
  if (!load_npy(dir + "/input.npy", "<c8", c, shape, 2)) return false;
  v.input.resize(c.size() / 2);
  std::memcpy(v.input.data(), c.data(), c.size() * sizeof(float));
  if (!load_npy(templates_file, "<c8", c, shape, 2)) return false;
  v.templates.resize(c.size() / 2);
  std::memcpy(v.templates.data(), c.data(), c.size() * sizeof(float));
  return load_npy(dir + "/thresholds.npy", "<f4", v.thresholds, shape) &&
         load_npy(dir + "/fop_ref.npy", "<f4", v.fop_ref, v.fop_shape) &&
         load_npy(dir + "/det_loc_ref.npy", "<u4", v.loc_ref, shape) &&
         load_npy(dir + "/det_pwr_ref.npy", "<f4", v.pwr_ref, shape);
}

}  // namespace

int main(int argc, char** argv) {
  std::vector<std::string> dirs;
  std::string templates_file;
  for (int i = 1; i < argc; ++i) {
    const std::string a = argv[i];
    if (a == "-N" && i + 1 < argc) g_runs = std::stoi(argv[++i]);
    else if (a == "-t" && i + 1 < argc) templates_file = argv[++i];
    else dirs.push_back(a);
  }
  if (dirs.empty() || templates_file.empty()) {
    std::fprintf(stderr, "usage: %s -t templates_ft_p4.npy [-N runs] vector_dir...\n", argv[0]);
    return 2;
  }
  for (const auto& d : dirs) {
    Vectors v;
    if (!load_vectors(d, templates_file, v)) {
      report("load " + d, false, "missing or malformed test vectors");
      continue;
    }
    std::printf("== test vectors %s (N=%zu)\n", d.c_str(), v.input.size());
    test_ft_convolution(v);
    test_harmonic_summing(v);
    test_steps(v);
    drive_serially(v, FDAS::PerSet, "FDAS_serial_set");
    drive_serially(v, FDAS::PerStage, "FDAS_serial_stage");
    drive_pipelined(v, FDAS::PerSet, "FDAS_pipelined_set");
    drive_pipelined(v, FDAS::PerStage, "FDAS_pipelined_stage");
    drive_pipelined(v, FDAS::Crossed, "FDAS_pipelined_crossed");
  }
  std::printf("%s: %d failure(s)\n", g_failures ? "FAILED" : "PASSED", g_failures);
  return g_failures ? 1 : 0;
}
