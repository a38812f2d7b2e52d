// Host-side planning shared by the single-trial context and the batch/serve engine.
#pragma once

#include <hip/hip_runtime.h>

#include <complex>
#include <cstdint>
#include <vector>

#include "fdas_jobs.h"
#include "fdas_kernels.h"

namespace fdas {

// Rows one FT-convolution work item convolves: enough to amortise the forward
// transform of its tile, small enough for several items per tile.  `target` is
// the preferred count; the rows are split into equal chunks of at most that.
uint32_t ftc_rows_per_item(uint32_t n_rows, uint32_t target = 16);

// Splits rows [row_begin, row_begin + n_rows) into contiguous groups, each with
// its own overlap (longest non-zero template of the group minus one, rounded up
// to 4).  row_taps may be empty (every row then uses max_taps).  The returned
// jobs have geometry and rows filled; pointers and item_begin are left to the
// caller.  The grouping (and so every FOP value) depends only on the tile, N and
// the tap extents; rows_target only sets how many rows one work item takes.
std::vector<FtcJob> plan_ftc_rows(uint32_t tile, uint32_t n_bins, uint32_t max_taps,
                                  const std::vector<uint32_t>& row_taps, uint32_t row_begin, uint32_t n_rows,
                                  uint32_t max_groups, uint32_t rows_target = 16);

// Template spectra (T x S, natural order, complex double) -> kernel layout
// [t][slot][tid], scaled by 1/S (exact).
std::vector<float2> kernel_template_layout(const std::vector<std::complex<double>>& spec, uint32_t n_templates,
                                           uint32_t tile);

// Taps (T x n_taps complex float) -> natural-order spectra of size tile; extent[t]
// receives the non-zero length of each row (at least 1).
std::vector<std::complex<double>> taps_to_spectra(const float* taps_c64, uint32_t n_templates, uint32_t n_taps,
                                                  uint32_t max_taps, uint32_t tile, std::vector<uint32_t>& extent);

// Twiddle tables of the S-point tile transform: w1[k][b] = W_S^(b k), w2[k][d] = W_(S/16)^(d k).
void make_twiddles(uint32_t tile, std::vector<float2>& tw1, std::vector<float2>& tw2);

// HSUM workgroup row chunking for a part of n_rows rows (at most `target` rows per workgroup).
void hsum_row_chunks(uint32_t n_rows, uint32_t* rows_per_block, uint32_t* n_row_chunks,
                     uint32_t target = kHsumRowsPerBlock);

}  // namespace fdas
