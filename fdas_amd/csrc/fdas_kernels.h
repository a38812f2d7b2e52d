// Host-side entry points of the device kernels (implemented in fdas_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "fdas_jobs.h"
#include "fdas_select.h"

namespace fdas {

constexpr int kHsumThreads = 256;
constexpr int kHsumBinsPerThread = 8;
constexpr int kHsumBinsPerBlock = kHsumThreads * kHsumBinsPerThread;  // bins one HSUM workgroup covers
// The 16-byte-load form of the summing kernel needs whole 8-bin groups.
inline bool hsum_vec_ok(uint32_t n_bins) { return n_bins % kHsumBinsPerThread == 0; }
constexpr uint32_t kHsumRowsPerBlock = 11;  // target rows per HSUM workgroup (bounds the dense-hit tail)
constexpr int kHsumWaveStage = 64;        // LDS staging entries per plane per wave before a global flush
constexpr int kSelectThreads = 1024;
constexpr uint32_t kSelectMaxCap = 8192;  // largest per-plane cap the in-LDS sort handles
constexpr uint32_t kSelectPool = 4096;    // boundary-bin entries resolved exactly in LDS
constexpr uint32_t kSelectSlices = 64;    // workgroups per plane in the histogram / compaction passes

int ftconv_threads(int tile);

// Fused FT convolution over a list of jobs (all with the same tile size).
hipError_t launch_ftconv(int tile, const FtcJob* d_jobs, uint32_t n_jobs, uint32_t n_items, const float2* tw1,
                         const float2* tw2, uint32_t grid, hipStream_t stream);

// Forward transform of every tile of one spectrum, natural bin order (debug / drop-in retrieve_tiles).
hipError_t launch_tiles(int tile, const float2* spectrum, uint32_t n_bins, uint32_t overlap, uint32_t payload,
                        uint32_t n_tiles, const float2* tw1, const float2* tw2, float2* out, hipStream_t stream);

// Harmonic summing + thresholding + compaction into per-plane spill lists.
// All jobs of one launch share n_harmonics; vec requires every job's n_bins % 4 == 0.
hipError_t launch_hsum(const HsumJob* d_jobs, uint32_t n_jobs, uint32_t n_blocks, uint32_t n_harmonics, bool vec,
                       hipStream_t stream);

// Per-plane cap (strongest first) and sort by (row, bin): power histogram and
// compaction over the spill lists (kSelectSlices workgroups per plane), then one
// workgroup per plane resolves the boundary bin exactly and sorts.
hipError_t launch_select(const SelectJob* d_jobs, uint32_t n_jobs, hipStream_t stream);

// Harmonic plane HP_k over rows [row_begin, row_begin + n_rows) (debug / parity).
hipError_t launch_hplane(const float* fop, uint32_t n_bins, uint32_t n_templates, uint32_t row_begin,
                         uint32_t n_rows, uint32_t k, float* out, hipStream_t stream);

// Per-block maxima of HP_1 .. HP_H over a part (blocks of 256 bins x 8 rows),
// out[k-1][block] for hmax_blocks(n_bins, n_rows) blocks per plane.
uint32_t hmax_blocks(uint32_t n_bins, uint32_t n_rows);
hipError_t launch_hmax(const float* fop, uint32_t n_bins, uint32_t n_templates, uint32_t row_begin, uint32_t n_rows,
                       uint32_t n_harmonics, float* out, hipStream_t stream);

hipError_t launch_zero_u32(uint32_t* p, uint32_t n, hipStream_t stream);

// One launch running both the FT convolution of ftc_jobs (1024-point tiles)
// and the harmonic summing of hsum_jobs (n_harmonics planes; vec as in
// launch_hsum) over a FOP written by an earlier launch; workgroups of the two
// kinds are interleaved.  fused_supported tells which shapes it covers.
bool fused_supported(uint32_t tile, uint32_t n_harmonics, bool vec);
hipError_t launch_fused(const FtcJob* ftc_jobs, uint32_t n_ftc_jobs, uint32_t n_items, const float2* tw1,
                        const float2* tw2, const HsumJob* hsum_jobs, uint32_t n_hsum_jobs, uint32_t n_blocks,
                        uint32_t n_harmonics, bool vec, hipStream_t stream);

}  // namespace fdas
