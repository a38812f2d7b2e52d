// Host-side test of the selection ordering rules (fdas_amd/csrc/fdas_select.h)
// that the selection kernels use: the order-preserving key, the histogram bin
// of a power, and the cut search split into per-thread shares exactly as the
// device's find_cut splits it.  Every case checks the cut against a brute-force
// sort; it includes planes whose detections are all negative or -0.0, which a
// caller can produce through inject_FOP and negative thresholds.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

#include "fdas_select.h"

using namespace fdas;

static int failures = 0;
#define CHECK(cond, ...)                        \
  do {                                          \
    if (!(cond)) {                              \
      std::printf("FAIL %s:%d: ", __FILE__, __LINE__); \
      std::printf(__VA_ARGS__);                 \
      std::printf("\n");                        \
      ++failures;                               \
    }                                           \
  } while (0)

static uint32_t bits(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  return u;
}

// find_cut as the device runs it with THREADS threads of PER bins each.
template <int THREADS>
static void cut_by_shares(const std::vector<uint32_t>& hist, uint32_t cap, uint32_t* bin, uint32_t* above) {
  constexpr int PER = int(kSelectBins) / THREADS;
  *bin = 0;
  *above = 0;
  uint32_t excl = 0;
  int found = 0;
  for (int t = 0; t < THREADS; ++t) {
    uint32_t h[PER];
    uint32_t local = 0;
    for (int j = 0; j < PER; ++j) {
      h[j] = hist[kSelectBins - 1 - uint32_t(t * PER + j)];
      local += h[j];
    }
    uint32_t b, a;
    if (cut_in_share<PER>(h, excl, cap, uint32_t(t * PER), &b, &a)) {
      *bin = b;
      *above = a;
      ++found;
    }
    excl += local;
  }
  CHECK(found == 1, "cut found by %d shares (want exactly 1)", found);
}

template <int THREADS>
static void check_plane(const char* name, const std::vector<float>& pwr, float thr, uint32_t cap) {
  const uint32_t base = histogram_base(bits(thr));
  std::vector<uint32_t> hist(kSelectBins, 0);
  for (float p : pwr) {
    const uint32_t b = power_bin(bits(p), base);
    CHECK(b < kSelectBins, "%s: bin %u out of range", name, b);
    hist[b] += 1;
  }
  uint32_t bstar, above;
  cut_by_shares<THREADS>(hist, cap, &bstar, &above);
  CHECK(bstar < kSelectBins, "%s: cut bin %u out of range", name, bstar);
  // brute force: the cap-th strongest power lies in the cut bin, the stronger bins hold `above` entries
  std::vector<float> s = pwr;
  std::sort(s.begin(), s.end(), [](float a, float b) { return a > b; });
  const uint32_t want = power_bin(bits(s[cap - 1]), base);
  uint32_t n_above = 0, n_at = 0;
  for (float p : pwr) {
    const uint32_t b = power_bin(bits(p), base);
    n_above += b > bstar;
    n_at += b == bstar;
  }
  CHECK(bstar == want, "%s: cut bin %u, brute force %u", name, bstar, want);
  CHECK(above == n_above && above < cap && cap <= above + n_at, "%s: above %u (brute %u), at %u, cap %u", name,
        above, n_above, n_at, cap);
}

int main() {
  // order_key orders floats like their values
  std::vector<float> vals = {-INFINITY, -1e30f, -2.5f, -1.0f, -1e-30f, -1e-45f, -0.0f, 0.0f, 1e-45f,
                             1e-30f,    0.5f,   1.0f,  3.0f,  1e30f,   INFINITY};
  for (size_t i = 0; i < vals.size(); ++i) {
    CHECK(order_key_inv(order_key(bits(vals[i]))) == bits(vals[i]), "order_key_inv round trip at %zu", i);
    for (size_t j = 0; j < vals.size(); ++j) {
      if (vals[i] > vals[j]) CHECK(order_key(bits(vals[i])) > order_key(bits(vals[j])), "order %zu > %zu", i, j);
      if (vals[i] < vals[j]) CHECK(order_key(bits(vals[i])) < order_key(bits(vals[j])), "order %zu < %zu", i, j);
    }
  }
  // power_bin is monotonic in the power, for any base
  for (float thr : {-3.0f, -0.0f, 0.0f, 1e-6f, 2.0f}) {
    const uint32_t base = histogram_base(bits(thr));
    for (size_t i = 0; i + 1 < vals.size(); ++i)
      CHECK(power_bin(bits(vals[i]), base) <= power_bin(bits(vals[i + 1]), base), "bin order at %zu thr %g", i,
            double(thr));
  }

  std::mt19937_64 rng(12345);
  const uint32_t cap = 4096;
  auto uniform = [&](float lo, float hi, size_t n) {
    std::uniform_real_distribution<float> u(lo, hi);
    std::vector<float> v(n);
    for (auto& x : v) x = u(rng);
    return v;
  };
  // all detections negative (threshold below them)
  check_plane<256>("all negative", uniform(-5.f, -1e-3f, 10000), -5.f, cap);
  check_plane<1024>("all negative (1024)", uniform(-5.f, -1e-3f, 10000), -5.f, cap);
  // all -0.0
  check_plane<256>("negative zeros", std::vector<float>(9000, -0.0f), -1.f, cap);
  // negative and positive detections
  check_plane<256>("mixed signs", uniform(-1.f, 100.f, 20000), -1.f, cap);
  check_plane<1024>("mixed signs (1024)", uniform(-1.f, 100.f, 20000), -1.f, cap);
  // the usual case: positive threshold, powers above it, including the top bin
  {
    std::vector<float> v = uniform(10.f, 1e4f, 30000);
    v.push_back(1e30f);
    check_plane<256>("positive", v, 10.f, cap);
  }
  // cap = number of entries - 1, and many equal powers
  check_plane<256>("ties", std::vector<float>(5000, 7.0f), 1.f, 4999);

  if (failures) {
    std::printf("%d failures\n", failures);
    return 1;
  }
  std::printf("select cut test ok\n");
  return 0;
}
