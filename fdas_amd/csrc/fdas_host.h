// Internal host-side declarations shared by the library's translation units.
#pragma once

#include <complex>
#include <cstdint>
#include <vector>

#include "../../include/fdas.h"

namespace fdas {

void fresnel(double x, double* s, double* c);
uint32_t drift_taps(double z);
std::vector<std::complex<double>> drift_response(double z);
void fft_inplace(std::complex<double>* a, uint32_t n, int dir);

// Frequency bin held by kernel thread `tid` in register slot `slot` after the
// forward tile transform of an S-point tile (matches TileFFT<S>::slot_bin).
inline uint32_t tile_slot_bin(uint32_t tile, uint32_t tid, uint32_t slot) {
  const uint32_t nt = tile / 16, d = nt / 16;
  return tid + nt * (slot / d) + 256u * (slot % d);
}

}  // namespace fdas
