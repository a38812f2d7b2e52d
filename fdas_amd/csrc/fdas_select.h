// Ordering rules of candidate selection, shared by the selection kernels and
// their host-side unit test.
//
// Selection keeps the `cap` strongest detections of a plane.  Its first pass
// bins every detection's power into a histogram (256 bins per octave above the
// plane's threshold); the bin holding the cap-th strongest entry is the cut.
// The rules must hold for any float a caller can inject (fdas_inject_fop takes
// arbitrary values and thresholds may be negative), so:
//  * every entry is counted: a power at or below the base, and every value with
//    the sign bit set (negatives, -0.0), lands in bin 0;
//  * exact comparisons use order_key, a 32-bit key that orders floats like the
//    values themselves (raw float bits misorder negatives).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace fdas {

constexpr uint32_t kSelectBins = 2048;  // power histogram bins (256 per octave above the threshold)

// a > b (as floats, no NaN)  <=>  order_key(a) > order_key(b).
__host__ __device__ inline uint32_t order_key(uint32_t bits) {
  return (bits & 0x80000000u) ? ~bits : (bits | 0x80000000u);
}
__host__ __device__ inline uint32_t order_key_inv(uint32_t key) {
  return (key & 0x80000000u) ? (key & 0x7fffffffu) : ~key;
}

// Histogram base of a plane: the threshold's bits when it is positive, else 0.
__host__ __device__ inline uint32_t histogram_base(uint32_t thr_bits) {
  return (thr_bits & 0x80000000u) || thr_bits == 0u ? 0u : thr_bits;
}

// Histogram bin of a power (float bits pbits); monotonic in the power.
__host__ __device__ inline uint32_t power_bin(uint32_t pbits, uint32_t base) {
  if ((pbits & 0x80000000u) || pbits <= base) return 0u;
  const uint32_t b = (pbits - base) >> 15;
  return b < kSelectBins - 1 ? b : kSelectBins - 1;
}

// One thread's share of the cut search.  The histogram is scanned from its top
// bin down; this thread owns PER consecutive bins of that order, first_from_top
// being the position of its first, and `excl` entries lie in the bins above
// its share.  If the cap-th entry from the top falls in the share, returns
// true with *bin = its histogram bin and *above = the entries in higher bins.
template <int PER>
__host__ __device__ inline bool cut_in_share(const uint32_t (&h)[PER], uint32_t excl, uint32_t cap,
                                             uint32_t first_from_top, uint32_t* bin, uint32_t* above) {
  uint32_t acc = excl;
  bool found = false;
  for (int j = 0; j < PER; ++j) {
    if (!found && acc < cap && cap <= acc + h[j]) {
      *bin = kSelectBins - 1 - (first_from_top + uint32_t(j));
      *above = acc;
      found = true;
    }
    acc += h[j];
  }
  return found;
}

}  // namespace fdas
