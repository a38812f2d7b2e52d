// Work descriptors passed from the host runtime to the device kernels.
//
// One descriptor describes one (trial, FOP part): the same structures serve the
// single-trial drop-in path (class FDAS, one job) and the batched / served
// paths (many jobs of uneven N, bank and harmonic count in one launch).
#pragma once

#include <cstdint>

#include "fdas_geometry.h"

namespace fdas {

struct FtcJob {
  const void* spectrum;   // float2[n_bins]
  float* fop;             // float[T * n_bins], row-major, row = t + c
  const void* hperm;      // float2[T][16][NT]: template spectra in kernel slot order, scaled 1/S
  uint32_t n_bins;
  uint32_t overlap;
  uint32_t payload;
  uint32_t n_tiles;
  uint32_t row_begin;     // first FOP row of the part
  uint32_t n_rows;        // rows of the part
  uint32_t rows_per_item; // rows one work item convolves
  uint32_t n_chunks;      // ceil(n_rows / rows_per_item)
  uint32_t item_begin;    // exclusive prefix of n_tiles * n_chunks over jobs
  uint32_t pad_;
};

struct HsumJob {
  const float* fop;       // float[T * n_bins]
  uint32_t n_bins;
  uint32_t n_templates;   // T (rows of the whole plane; c = T / 2)
  uint32_t row_begin;
  uint32_t n_rows;
  uint32_t n_harmonics;
  uint32_t spill_cap;     // entries per plane in the spill lists
  uint32_t block_begin;   // exclusive prefix of workgroups over jobs
  uint32_t rows_per_block;  // rows one workgroup sums (the part is split into row chunks)
  uint32_t n_row_chunks;    // ceil(n_rows / rows_per_block); workgroups = bin blocks * n_row_chunks
  uint32_t pad_;
  uint32_t* spill_key;    // [H][spill_cap] row << 22 | bin
  float* spill_pwr;       // [H][spill_cap]
  uint32_t* counts;       // [H] detections per plane (may exceed spill_cap)
  float thr[kMaxHarmonics];
};

struct SelectJob {
  const uint32_t* spill_key;  // [spill_cap]
  const float* spill_pwr;
  const uint32_t* count;      // detections of this plane
  uint32_t spill_cap;
  uint32_t cap;               // entries kept
  float thr;                  // the plane's threshold: base of the power histogram
  uint32_t compact_cap;       // capacity of the compact list (cap + boundary pool)
  uint32_t* hist;             // [kSelectBins] power histogram of the spill (zeroed per run)
  uint32_t* compact_key;      // [compact_cap] entries at or above the cut bin
  float* compact_pwr;
  uint32_t* compact_n;        // zeroed per run
  uint32_t* out_key;          // [cap] sorted by key
  float* out_pwr;
  uint32_t* out_n;            // kept entries
};

}  // namespace fdas
