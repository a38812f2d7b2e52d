// Geometry and candidate encoding shared by host and device code.
//
// Follows the reference's compile-time GenInfo (fdas/src/host/gen_info.h:26-74)
// and host geometry (fdas/src/host/FDAS.cpp:177-186), made run-time values so
// that one library serves every N, bank size, tile size and harmonic count.
#pragma once

#include <cstdint>

namespace fdas {

constexpr uint32_t kPack = 4;              // overlap rounding unit, cl_codegen.py:99
constexpr uint32_t kMaxHarmonics = 16;     // widest plane count supported by the kernels
constexpr uint32_t kMaxBins = 1u << 22;    // 22-bit frequency field, gen_info.h:62-67
constexpr uint32_t kMaxTemplates = 255;    // 8-bit row field of the internal candidate key
constexpr uint32_t kKeyBinBits = 22;       // internal key = row << 22 | bin

// Overlap O = ceil((L_max - 1) / 4) * 4  (cl_codegen.py:99-100, ft_conv.py:161).
inline constexpr uint32_t overlap_for(uint32_t max_taps) {
  return max_taps == 0 ? 0 : ((max_taps - 1 + kPack - 1) / kPack) * kPack;
}

struct Geometry {
  uint32_t n_bins = 0;
  uint32_t tile = 0;
  uint32_t overlap = 0;
  uint32_t payload = 0;
  uint32_t n_tiles = 0;
  uint32_t padding_last_tile = 0;
};

// n_tiles = ceil(N / P) and the zero padding of the last tile (FDAS.cpp:177-184).
inline Geometry make_geometry(uint32_t n_bins, uint32_t tile, uint32_t max_taps) {
  Geometry g;
  g.n_bins = n_bins;
  g.tile = tile;
  g.overlap = overlap_for(max_taps);
  g.payload = tile > g.overlap ? tile - g.overlap : 0;
  if (g.payload == 0) return g;
  g.n_tiles = n_bins / g.payload + 1;
  g.padding_last_tile = g.payload - n_bins % g.payload;
  if (g.padding_last_tile == g.payload) {
    --g.n_tiles;
    g.padding_last_tile = 0;
  }
  return g;
}

// Row of FOP[trunc(t / k)] for row r = t + c: truncation toward zero
// (hsum.py:109-114, hms_baseline.cl.mako:66-68).
inline constexpr int sub_row(int r, int k, int c) { return c + (r - c) / k; }

// Default spill capacity per plane: at least 262,144 entries and T*N/512, so
// noise alone at the (1 + 6) x mean-power plane-1 threshold (an exp(-7)
// ~ 1/1100 fraction of the plane) stays within it.
inline constexpr uint32_t default_spill(uint32_t n_templates, uint32_t n_bins) {
  const uint64_t by_size = (uint64_t(n_templates) * n_bins + 511) / 512;
  return uint32_t(by_size > 262144u ? by_size : 262144u);
}

// 32-bit candidate location of the reference (gen_info.h:62-69, hsum.py:147).
inline constexpr uint32_t encode_location(uint32_t harm, int tmpl, uint32_t freq, int c) {
  return (((harm - 1) & 0x7u) << 29) | ((uint32_t(tmpl + c) & 0x7fu) << 22) | (freq & 0x3fffffu);
}
inline constexpr uint32_t get_harmonic(uint32_t loc) { return ((loc >> 29) & 0x7u) + 1; }
inline constexpr int get_template_num(uint32_t loc, int c) { return int((loc >> 22) & 0x7fu) - c; }
inline constexpr uint32_t get_frequency_bin(uint32_t loc) { return loc & 0x3fffffu; }
inline constexpr uint32_t invalid_location(int c) { return encode_location(1, c + 1, 0, c); }

static_assert(invalid_location(42) == 0x15400000u, "reference invalid slot, gen_info.h:69");

}  // namespace fdas
