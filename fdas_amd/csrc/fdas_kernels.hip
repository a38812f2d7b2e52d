// Device kernels of the FDAS hot path for CDNA4 (gfx950).
//
//  ftconv_kernel  fused overlap-save FT convolution: tile load -> forward FFT
//                 -> x template spectrum -> inverse FFT -> discard overlap -> |x|^2
//                 (reference: tile_input.cl.mako, fft.cl.mako, mux_and_mult.cl.mako,
//                  square_and_discard.cl.mako; oracle ft_conv.py:158-207)
//  hsum_kernel    harmonic summing + per-plane strict thresholding + compaction
//                 (reference: hms_baseline.cl.mako:55-71, detect.cl.mako:81-104,
//                  hsum.py:91-151)
//  select_kernel  per-plane cap (strongest first) and (row, bin) ordering of the
//                 detections; replaces the reference's overwrite-oldest ring
//                 (detect.cl.mako:106-115) with a deterministic rule.
#include <hip/hip_runtime.h>

#include <climits>
#include <type_traits>

#include "fdas_kernels.h"
#include "fft_tile.h"

namespace fdas {

// ---------------------------------------------------------------------------
// FT convolution
// ---------------------------------------------------------------------------

__device__ __forceinline__ uint32_t find_job_ftc(const FtcJob* __restrict__ jobs, uint32_t n_jobs, uint32_t item) {
  uint32_t lo = 0, hi = n_jobs;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (jobs[mid].item_begin <= item) lo = mid; else hi = mid;
  }
  return lo;
}

template <int S>
__global__ __launch_bounds__(S / 16) void ftconv_kernel(const FtcJob* __restrict__ jobs, uint32_t n_jobs,
                                                        uint32_t n_items, const float2* __restrict__ tw1,
                                                        const float2* __restrict__ tw2) {
  using F = TileFFT<S>;
  constexpr int NT = F::NT;
  // A one-wave workgroup (S = 1024) barriers for free: it keeps one LDS buffer
  // (8.7 KB), so more workgroups fit beside it on a CU.  Wider tiles keep two
  // buffers and save two barriers per inverse transform.
  constexpr bool kSingle = (S == 1024);
  __shared__ float2 la[F::LDS_ELEMS];
  __shared__ float2 lb[kSingle ? 1 : F::LDS_ELEMS];
  const int tid = threadIdx.x;
  F fft;
  fft.load_twiddles(tw1, tw2, tid);

  for (uint32_t item = blockIdx.x; item < n_items; item += gridDim.x) {
    const FtcJob* job = jobs + find_job_ftc(jobs, n_jobs, item);
    const uint32_t n_bins = job->n_bins;
    const uint32_t overlap = job->overlap;
    const uint32_t payload = job->payload;
    const uint32_t local = item - job->item_begin;
    const uint32_t chunk = local / job->n_tiles;  // consecutive items share a template chunk
    const uint32_t tile = local - chunk * job->n_tiles;
    const float2* __restrict__ x = static_cast<const float2*>(job->spectrum);

    __syncthreads();  // previous item's LDS readers are done

    float2 xf[16];
    const int base = int(tile * payload) - int(overlap);
#pragma unroll
    for (int a = 0; a < 16; ++a) {
      const int idx = base + NT * a + tid;
      xf[a] = make_float2(0.f, 0.f);
      if (idx >= 0 && idx < int(n_bins)) xf[a] = gld2(x, size_t(idx));
    }

    const uint32_t r_begin = job->row_begin + chunk * job->rows_per_item;
    uint32_t r_end = r_begin + job->rows_per_item;
    const uint32_t r_stop = job->row_begin + job->n_rows;
    if (r_end > r_stop) r_end = r_stop;
    const float2* __restrict__ hbase = static_cast<const float2*>(job->hperm);
    const uint32_t out0 = tile * payload;  // FOP bin of tile sample s = overlap

    // Template spectra are loaded one row ahead: the first row's before the
    // forward transform, every next row's as soon as this row's has been used,
    // so the loads have a whole transform to complete (and they precede this
    // row's FOP stores: loads and stores share one in-order completion counter,
    // so a load issued after the stores would wait for them).
    float2 hn[16];
#pragma unroll
    for (int s = 0; s < 16; ++s) hn[s] = gld2(hbase + size_t(r_begin) * S + tid, size_t(s * NT));
    fft.forward(xf, la, tid);
    for (uint32_t r = r_begin; r < r_end; ++r) {
      float2 v[16];
#pragma unroll
      for (int s = 0; s < 16; ++s) v[s] = cmul(xf[s], hn[s]);
      {
        const float2* __restrict__ h = hbase + size_t(min(r + 1, r_end - 1)) * S + tid;  // unconditional: no branch
#pragma unroll
        for (int s = 0; s < 16; ++s) hn[s] = gld2(h, size_t(s * NT));
      }
      if constexpr (kSingle)
        fft.inverse_single(v, la, tid);
      else
        fft.inverse(v, la, lb, tid);
      // Slot a of this thread is FOP bin out0 + NT*a + tid - overlap: one base
      // address per row and a constant offset per slot (the store's immediate).
      float* __restrict__ out = job->fop + size_t(r) * n_bins + (ptrdiff_t(out0) + tid - ptrdiff_t(overlap));
#pragma unroll
      for (int a = 0; a < 16; ++a) {
        const uint32_t s = uint32_t(NT * a + tid);
        if (s >= overlap) {
          const uint32_t f = out0 + s - overlap;
          if (f < n_bins) gst_stream(out, size_t(NT * a), v[a].x * v[a].x + v[a].y * v[a].y);
        }
      }
    }
  }
}

// 1024-point work items for four-wave workgroups (the fused launch below):
// wave 0 loads the tile and forward-transforms it into LDS (xs); then every
// wave convolves its share of the item's template rows (rows w, w + 4, ...),
// each row reading the tile spectrum back from LDS, so the workgroup runs at
// three waves per SIMD.
constexpr int kFtc1kWaves = 4;

// LDS of one 1024-point work item, carved from a caller's buffer.
struct Ftc1kLds {
  float2 xs[16 * WaveFFT1k<true>::NT];                   // the tile's spectrum, slot-major
  float2 lbuf[kFtc1kWaves][WaveFFT1k<true>::LDS_ELEMS];  // each wave's exchange buffer
  float2 w2s[64];                                  // W_64^(d * kc), the F2 twiddles
};

// One work item (tile, chunk of template rows) of a 1024-point job list.
// `fft` holds this thread's twiddles; `first` is true for the workgroup's
// first item (wave 0 has just filled w2s: the first barrier publishes it).
__device__ __forceinline__ void ftc1k_item(Ftc1kLds& L, const WaveFFT1k<true>& fft,
                                           const FtcJob* __restrict__ jobs, uint32_t n_jobs, uint32_t item) {
  using F = WaveFFT1k<true>;
  constexpr int NT = F::NT;
  constexpr int S = F::S;
  const int lane = int(threadIdx.x) & 63;
  const uint32_t wave = threadIdx.x >> 6;
  float2* __restrict__ lb = L.lbuf[wave];
  const FtcJob* job = jobs + find_job_ftc(jobs, n_jobs, item);
  const uint32_t n_bins = job->n_bins;
  const uint32_t overlap = job->overlap;
  const uint32_t payload = job->payload;
  const uint32_t local = item - job->item_begin;
  const uint32_t chunk = local / job->n_tiles;  // consecutive items share a template chunk
  const uint32_t tile = local - chunk * job->n_tiles;
  const uint32_t r_begin = job->row_begin + chunk * job->rows_per_item;
  uint32_t r_end = r_begin + job->rows_per_item;
  const uint32_t r_stop = job->row_begin + job->n_rows;
  if (r_end > r_stop) r_end = r_stop;
  const float2* __restrict__ hbase = static_cast<const float2*>(job->hperm);
  const uint32_t out0 = tile * payload;  // FOP bin of tile sample s = overlap

  // This wave's first template row is loaded before the forward transform,
  // every next one as soon as the current one has been used.
  uint32_t r = r_begin + wave;
  float2 hn[16];
  {
    const float2* __restrict__ h = hbase + size_t(min(r, r_end - 1)) * S + lane;
#pragma unroll
    for (int s = 0; s < 16; ++s) hn[s] = gld2(h, size_t(s * NT));
  }

  __syncthreads();  // every wave has finished reading the previous item's spectrum
  if (wave == 0) {
    const float2* __restrict__ x = static_cast<const float2*>(job->spectrum);
    float2 v[16];
    const int base = int(tile * payload) - int(overlap);
#pragma unroll
    for (int a = 0; a < 16; ++a) {
      const int idx = base + NT * a + lane;
      v[a] = make_float2(0.f, 0.f);
      if (idx >= 0 && idx < int(n_bins)) v[a] = gld2(x, size_t(idx));
    }
    fft.forward(v, lb, lane);
#pragma unroll
    for (int s = 0; s < 16; ++s) L.xs[s * NT + lane] = v[s];
  }
  __syncthreads();

  for (; r < r_end; r += kFtc1kWaves) {
    float2 v[16];
#pragma unroll
    for (int s = 0; s < 16; ++s) v[s] = cmul(L.xs[s * NT + lane], hn[s]);
    {
      const float2* __restrict__ h = hbase + size_t(min(r + kFtc1kWaves, r_end - 1)) * S + lane;
#pragma unroll
      for (int s = 0; s < 16; ++s) hn[s] = gld2(h, size_t(s * NT));
    }
    fft.inverse(v, lb, lane);
    float* __restrict__ out = job->fop + size_t(r) * n_bins + (ptrdiff_t(out0) + lane - ptrdiff_t(overlap));
#pragma unroll
    for (int a = 0; a < 16; ++a) {
      const uint32_t s = uint32_t(NT * a + lane);
      if (s >= overlap) {
        const uint32_t f = out0 + s - overlap;
        if (f < n_bins) gst_stream(out, size_t(NT * a), v[a].x * v[a].x + v[a].y * v[a].y);
      }
    }
  }
}

// 1024-point tiles, one wave per workgroup.  The tile's spectrum and the
// twiddles stay in registers; the F2 <-> F3 exchange is a pair of permlane
// swaps and only the F1 <-> F2 exchange goes through LDS (8.3 KB per
// workgroup).  164 VGPRs: three waves per SIMD.
__global__ __launch_bounds__(64) void ftconv1w_kernel(const FtcJob* __restrict__ jobs, uint32_t n_jobs,
                                                      uint32_t n_items, const float2* __restrict__ tw1,
                                                      const float2* __restrict__ tw2) {
  using F = WaveFFT1kLean;
  constexpr int NT = F::NT;
  constexpr int S = F::S;
  __shared__ float2 lb[F::LDS_ELEMS];
  const int lane = threadIdx.x;
  F fft;
  fft.load_twiddles(tw1, tw2, lane);

  for (uint32_t item = blockIdx.x; item < n_items; item += gridDim.x) {
    const FtcJob* job = jobs + find_job_ftc(jobs, n_jobs, item);
    const uint32_t n_bins = job->n_bins;
    const uint32_t overlap = job->overlap;
    const uint32_t payload = job->payload;
    const uint32_t local = item - job->item_begin;
    const uint32_t chunk = local / job->n_tiles;
    const uint32_t tile = local - chunk * job->n_tiles;
    const float2* __restrict__ x = static_cast<const float2*>(job->spectrum);
    const uint32_t r_begin = job->row_begin + chunk * job->rows_per_item;
    uint32_t r_end = r_begin + job->rows_per_item;
    const uint32_t r_stop = job->row_begin + job->n_rows;
    if (r_end > r_stop) r_end = r_stop;
    const float2* __restrict__ hbase = static_cast<const float2*>(job->hperm);
    const uint32_t out0 = tile * payload;

    float2 xf[16];
    const int base = int(tile * payload) - int(overlap);
#pragma unroll
    for (int a = 0; a < 16; ++a) {
      const int idx = base + NT * a + lane;
      xf[a] = make_float2(0.f, 0.f);
      if (idx >= 0 && idx < int(n_bins)) xf[a] = gld2(x, size_t(idx));
    }
    fft.forward(xf, lb, lane);
    for (uint32_t r = r_begin; r < r_end; ++r) {
      // The template spectrum is not prefetched: at three waves per SIMD the
      // other waves' transforms cover the load.
      float2 v[16];
#pragma unroll
      for (int s = 0; s < 16; ++s) v[s] = cmul(xf[s], gld2(hbase + size_t(r) * S + lane, size_t(s * NT)));
      fft.inverse(v, lb, lane);
      float* __restrict__ out = job->fop + size_t(r) * n_bins + (ptrdiff_t(out0) + lane - ptrdiff_t(overlap));
#pragma unroll
      for (int a = 0; a < 16; ++a) {
        const uint32_t s = uint32_t(NT * a + lane);
        if (s >= overlap) {
          const uint32_t f = out0 + s - overlap;
          if (f < n_bins) gst_stream(out, size_t(NT * a), v[a].x * v[a].x + v[a].y * v[a].y);
        }
      }
    }
  }
}

template <int S>
__global__ __launch_bounds__(S / 16) void tiles_kernel(const float2* __restrict__ x, uint32_t n_bins,
                                                       uint32_t overlap, uint32_t payload,
                                                       const float2* __restrict__ tw1,
                                                       const float2* __restrict__ tw2, float2* __restrict__ out) {
  using F = TileFFT<S>;
  constexpr int NT = F::NT;
  __shared__ float2 la[F::LDS_ELEMS];
  const int tid = threadIdx.x;
  F fft;
  fft.load_twiddles(tw1, tw2, tid);
  const uint32_t tile = blockIdx.x;
  const int base = int(tile * payload) - int(overlap);
  float2 v[16];
#pragma unroll
  for (int a = 0; a < 16; ++a) {
    const int idx = base + NT * a + tid;
    v[a] = (idx >= 0 && idx < int(n_bins)) ? x[idx] : make_float2(0.f, 0.f);
  }
  fft.forward(v, la, tid);
  float2* __restrict__ o = out + size_t(tile) * S;
#pragma unroll
  for (int s = 0; s < 16; ++s) o[F::slot_bin(tid, s)] = v[s];
}

int ftconv_threads(int tile) { return tile == 1024 ? 64 : tile / 16; }

hipError_t launch_ftconv(int tile, const FtcJob* d_jobs, uint32_t n_jobs, uint32_t n_items, const float2* tw1,
                         const float2* tw2, uint32_t grid, hipStream_t stream) {
  if (n_items == 0) return hipSuccess;
  if (grid == 0 || grid > n_items) grid = n_items;
  switch (tile) {
    case 1024: hipLaunchKernelGGL(ftconv1w_kernel, dim3(grid), dim3(64), 0, stream, d_jobs, n_jobs, n_items, tw1, tw2); break;
    case 2048: hipLaunchKernelGGL(ftconv_kernel<2048>, dim3(grid), dim3(128), 0, stream, d_jobs, n_jobs, n_items, tw1, tw2); break;
    case 4096: hipLaunchKernelGGL(ftconv_kernel<4096>, dim3(grid), dim3(256), 0, stream, d_jobs, n_jobs, n_items, tw1, tw2); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_tiles(int tile, const float2* spectrum, uint32_t n_bins, uint32_t overlap, uint32_t payload,
                        uint32_t n_tiles, const float2* tw1, const float2* tw2, float2* out, hipStream_t stream) {
  if (n_tiles == 0) return hipSuccess;
  switch (tile) {
    case 1024: hipLaunchKernelGGL(tiles_kernel<1024>, dim3(n_tiles), dim3(64), 0, stream, spectrum, n_bins, overlap, payload, tw1, tw2, out); break;
    case 2048: hipLaunchKernelGGL(tiles_kernel<2048>, dim3(n_tiles), dim3(128), 0, stream, spectrum, n_bins, overlap, payload, tw1, tw2, out); break;
    case 4096: hipLaunchKernelGGL(tiles_kernel<4096>, dim3(n_tiles), dim3(256), 0, stream, spectrum, n_bins, overlap, payload, tw1, tw2, out); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Harmonic summing
// ---------------------------------------------------------------------------

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

__device__ __forceinline__ uint32_t lanes_below(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi(uint32_t(mask >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(mask), 0u));
}

// Exclusive prefix sum over the wave of a per-lane count v < 16, bit-sliced:
// one ballot per bit.  `total` receives the wave's sum (wave-uniform).
__device__ __forceinline__ uint32_t wave_prefix_small(uint32_t v, uint32_t& total) {
  uint32_t excl = 0;
  total = 0;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const uint64_t m = __ballot((v >> j) & 1u);
    excl += lanes_below(m) << j;
    total += uint32_t(__builtin_popcountll(m)) << j;
  }
  return excl;
}

// Per-wave detection staging: each wave owns kHsumWaveStage LDS slots per
// plane and a wave-uniform fill count.  A full stage is flushed by the wave
// alone (one global atomic for up to 64 entries); what is left at the end is
// flushed with one atomic per plane per workgroup.
template <int H>
struct HsumStage {
  static constexpr int kCap = kHsumWaveStage;
  uint32_t key[kHsumThreads / 64][H][kCap];
  float pwr[kHsumThreads / 64][H][kCap];
  uint32_t n[kHsumThreads / 64][H];
  uint32_t base[kHsumThreads / 64][H];
};

// Writes the wave's n staged entries of plane p to the global spill list.
template <class Stage>
__device__ __forceinline__ void wave_flush(Stage& st, const HsumJob& job, uint32_t wave, uint32_t p, uint32_t n) {
  const uint32_t lane = lane_id();
  uint32_t g = 0;
  if (lane == 0) g = gatomic_add(job.counts + p, n);
  g = __shfl(g, 0);
  if (lane < n && g + lane < job.spill_cap) {
    gst(job.spill_key, size_t(p) * job.spill_cap + g + lane, st.key[wave][p][lane]);
    gst(job.spill_pwr, size_t(p) * job.spill_cap + g + lane, st.pwr[wave][p][lane]);
  }
}

__device__ __forceinline__ uint32_t find_job_hsum(const HsumJob* __restrict__ jobs, uint32_t n_jobs, uint32_t blk) {
  uint32_t lo = 0, hi = n_jobs;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (jobs[mid].block_begin <= blk) lo = mid; else hi = mid;
  }
  return lo;
}

// A thread sums 8 consecutive bins f0 .. f0+7 of a row, f0 a multiple of 8.
// For harmonic k their sub-harmonic bins (f0 + b) / k are the hsum_span(k)
// consecutive bins from f0 / k: f0 mod k is a multiple of g = gcd(8, k), at
// most k - g, so the last one is at most (k - g + 7) / k bins further on.
__host__ __device__ constexpr int hsum_gcd8(int k) { return (k % 8 == 0) ? 8 : (k % 4 == 0) ? 4 : (k % 2 == 0) ? 2 : 1; }
__host__ __device__ constexpr int hsum_span(int k) { return 1 + (k - hsum_gcd8(k) + 7) / k; }
constexpr int kHsumMaxSpan = 4;  // hsum_span(2) = hsum_span(3) = 4 is the largest
static_assert(kHsumBinsPerThread == 8, "the sub-harmonic spans assume 8 bins per thread");

// Every FOP value one thread needs for one row: its 8 own bins and, for each
// harmonic k >= 2, the hsum_span(k) sub-harmonic bins of row trunc(t / k).
template <int H>
struct RowLoads {
  float m[kHsumBinsPerThread];
  float s[H][kHsumMaxSpan];
};

// Branch-free: H is a compile-time constant and the indexes are clamped, so
// the compiler can count the loads in flight exactly (a load issued under a
// condition makes it drain every outstanding load at the next use).
//  VEC (n_bins % 8 == 0): f0 is clamped to the last 8 bins, so every index is
//    in range; the own bins are two 16-byte loads, and so are the k = 2 bins
//    (f0 / 2 is a multiple of 4); the k = 4 pair is one 8-byte load.
//  otherwise: one load per value, each index clamped to the last bin, and the
//    own bins past the end read as -inf, so no sum of theirs passes a threshold.
template <bool VEC>
__device__ __forceinline__ void hsum_load_own(float (&m)[kHsumBinsPerThread], const float* __restrict__ fop,
                                              uint32_t n_bins, uint32_t r, uint32_t f0) {
  const float* __restrict__ row = fop + size_t(r) * n_bins;
  const uint32_t last = n_bins - 1;
  if constexpr (VEC) {
    const fdas_vf4 q0 = gld4(row, f0);
    const fdas_vf4 q1 = gld4(row, f0 + 4);
    m[0] = q0.x; m[1] = q0.y; m[2] = q0.z; m[3] = q0.w;
    m[4] = q1.x; m[5] = q1.y; m[6] = q1.z; m[7] = q1.w;
  } else {
#pragma unroll
    for (int b = 0; b < kHsumBinsPerThread; ++b) {
      const float v = gld(row, min(f0 + uint32_t(b), last));
      m[b] = (f0 + uint32_t(b) < n_bins) ? v : -INFINITY;
    }
  }
}

// The sub-harmonic spans of row r (harmonics 2..H) for the 8 bins from f0.
template <int H, bool VEC>
__device__ __forceinline__ void hsum_load_subs(RowLoads<H>& L, const float* __restrict__ fop, uint32_t n_bins, int c,
                                               uint32_t r, uint32_t f0) {
  const uint32_t last = n_bins - 1;
  const int t = int(r) - c;
#pragma unroll
  for (int k = 2; k <= H; ++k) {
    const float* __restrict__ srow = fop + size_t(c + t / k) * n_bins;
    const uint32_t s0 = f0 / uint32_t(k);
    if (VEC && k == 2) {
      const fdas_vf4 q = gld4(srow, s0);
      L.s[k - 1][0] = q.x; L.s[k - 1][1] = q.y; L.s[k - 1][2] = q.z; L.s[k - 1][3] = q.w;
    } else if (VEC && k == 4) {
      const float2 q = gld2(reinterpret_cast<const float2*>(srow), s0 / 2);
      L.s[k - 1][0] = q.x; L.s[k - 1][1] = q.y;
    } else {
#pragma unroll
      for (int j = 0; j < hsum_span(k); ++j) L.s[k - 1][j] = gld(srow, VEC ? s0 + uint32_t(j) : min(s0 + uint32_t(j), last));
    }
  }
}

template <int H, bool VEC>
__device__ __forceinline__ void hsum_load_row(RowLoads<H>& L, const float* __restrict__ fop, uint32_t n_bins, int c,
                                              uint32_t r, uint32_t f0) {
  hsum_load_own<VEC>(L.m, fop, n_bins, r, f0);
  hsum_load_subs<H, VEC>(L, fop, n_bins, c, r, f0);
}

// Bin b of the thread takes sub-harmonic value (f0 mod k + b) / k of its span:
// b / k when f0 mod k < k (b / k + 1) - b, else the next one.  rk[k-1] = f0 mod k.
template <int K>
__device__ __forceinline__ float hsum_sub(const float (&s)[kHsumMaxSpan], uint32_t rk, int b) {
  const int lo = b / K;
  const int edge = K * (lo + 1) - b;  // f0 mod k at which bin b moves to the next value
  if (edge > K - hsum_gcd8(K)) return s[lo];  // no f0 mod k reaches it
  return rk >= uint32_t(edge) ? s[lo + 1] : s[lo];
}

template <int H, int K>
__device__ __forceinline__ void hsum_add(float (&hp)[kHsumBinsPerThread], const RowLoads<H>& L, uint32_t rk) {
#pragma unroll
  for (int b = 0; b < kHsumBinsPerThread; ++b) hp[b] += hsum_sub<K>(L.s[K - 1], rk, b);
}

template <int H, int K = 2>
__device__ __forceinline__ void hsum_add_k(int k, float (&hp)[kHsumBinsPerThread], const RowLoads<H>& L,
                                           const uint32_t (&rk)[H]) {
  if constexpr (K <= H) {
    if (k == K) hsum_add<H, K>(hp, L, rk[K - 1]);
    else hsum_add_k<H, K + 1>(k, hp, L, rk);
  }
}

// Key of the per-plane group test: a value v > thr implies
// float_as_int(v) >= group_key(thr) for every v and thr (a non-negative thr
// orders like its bit pattern against any v above it; a negative one admits
// everything), so the test below only ever lets through a superset of the
// bins the exact float compare accepts.
__device__ __forceinline__ int group_key(float thr) { return thr < 0.f ? INT_MIN : __float_as_int(thr); }

// Largest bit pattern of the 8 sums as signed integers: three max3 and one max,
// with no NaN canonicalisation of the operands (which fmaxf would need).
__device__ __forceinline__ int group_max(const float (&hp)[kHsumBinsPerThread]) {
  const int a = max(max(__float_as_int(hp[0]), __float_as_int(hp[1])), __float_as_int(hp[2]));
  const int b = max(max(__float_as_int(hp[3]), __float_as_int(hp[4])), __float_as_int(hp[5]));
  return max(max(a, b), max(__float_as_int(hp[6]), __float_as_int(hp[7])));
}

// One test per plane for the common case: the group key of the 8 sums against
// the plane's; only a wave with a candidate looks at the bins one by one with
// the exact float compare.  The thresholds are wave-uniform; `live_lanes`
// holds the lanes with at least one bin in range (per-bin range checks are in
// `live_bits`, bit b for bin b).
template <int H, class Stage>
__device__ __forceinline__ void hsum_row(Stage& st, const HsumJob& job, const RowLoads<H>& L, uint32_t r,
                                         uint32_t f0, const uint32_t (&rk)[H], const float (&thr)[H],
                                         const int (&thr_key)[H], uint64_t live_lanes, uint32_t live_bits,
                                         uint32_t (&wn)[H], uint32_t wave) {
  float hp[kHsumBinsPerThread];
#pragma unroll
  for (int b = 0; b < kHsumBinsPerThread; ++b) hp[b] = L.m[b];
#pragma unroll
  for (int k = 1; k <= H; ++k) {
    if (k > 1) hsum_add_k<H>(k, hp, L, rk);
    if ((__ballot(group_max(hp) >= thr_key[k - 1]) & live_lanes) == 0) continue;
    // The wave's hits in lane order, each lane's in bin order: one prefix sum
    // over the per-lane counts places them all.
    uint32_t bits = 0;
#pragma unroll
    for (int b = 0; b < kHsumBinsPerThread; ++b) bits |= (hp[b] > thr[k - 1]) ? (1u << b) : 0u;
    bits &= live_bits;
    uint32_t total;
    uint32_t pos = wave_prefix_small(uint32_t(__builtin_popcount(bits)), total);
    if (wn[k - 1] + total > uint32_t(Stage::kCap)) {
      wave_flush(st, job, wave, uint32_t(k - 1), wn[k - 1]);
      wn[k - 1] = 0;
    }
    if (total > uint32_t(Stage::kCap)) {
      // more than a stage holds: straight to the spill list (entries past its end are counted, not kept)
      uint32_t g = 0;
      if (lane_id() == 0) g = gatomic_add(job.counts + (k - 1), total);
      g = __shfl(g, 0) + pos;
#pragma unroll
      for (int b = 0; b < kHsumBinsPerThread; ++b) {
        if ((bits >> b) & 1u) {
          if (g < job.spill_cap) {
            gst(job.spill_key, size_t(k - 1) * job.spill_cap + g, (r << kKeyBinBits) | (f0 + uint32_t(b)));
            gst(job.spill_pwr, size_t(k - 1) * job.spill_cap + g, hp[b]);
          }
          ++g;
        }
      }
    } else {
      pos += wn[k - 1];
#pragma unroll
      for (int b = 0; b < kHsumBinsPerThread; ++b) {
        if ((bits >> b) & 1u) {
          st.key[wave][k - 1][pos] = (r << kKeyBinBits) | (f0 + uint32_t(b));
          st.pwr[wave][k - 1][pos] = hp[b];
          ++pos;
        }
      }
      wn[k - 1] += total;
    }
  }
}

// HP_1 = FOP, HP_k = HP_{k-1} + FOP[trunc(t/k)][f/k] in FP32, in k order (hsum.py:93-120);
// strict HP_k > thr_k (detect.cl.mako:81-83).  One workgroup covers 2048 bins
// of a chunk of rows of its part, each thread 8 consecutive bins.  Rows are
// processed in pairs with two register buffers: the loads of row r+1 are in
// flight while row r is summed and thresholded.
template <int H, bool VEC>
__device__ __forceinline__ void hsum_block(HsumStage<H>& st, const HsumJob* __restrict__ jobs, uint32_t n_jobs,
                                           uint32_t blk) {
  const HsumJob& job = jobs[find_job_hsum(jobs, n_jobs, blk)];
  const int tid = threadIdx.x;
  const uint32_t wave = uint32_t(tid) >> 6;

  const uint32_t n_bins = job.n_bins;
  const int c = int(job.n_templates / 2);
  const float* __restrict__ fop = job.fop;
  const uint32_t local = blk - job.block_begin;
  const uint32_t chunk = local % job.n_row_chunks;  // row chunks of one bin block run side by side
  const uint32_t f0 = (local / job.n_row_chunks) * uint32_t(kHsumBinsPerBlock) + uint32_t(tid * kHsumBinsPerThread);
  // VEC: threads past the end read (and discard) the last 8 bins; both values are multiples of 8
  const uint32_t fl = VEC ? min(f0, (n_bins - 8u) & ~7u) : f0;
  const uint32_t r_begin = job.row_begin + chunk * job.rows_per_block;
  const uint32_t r_last = min(r_begin + job.rows_per_block, job.row_begin + job.n_rows) - 1;
  float thr[H];
  int thr_key[H];
  uint32_t wn[H], rk[H];
#pragma unroll
  for (int k = 0; k < H; ++k) {
    thr[k] = job.thr[k];
    thr_key[k] = group_key(thr[k]);
    wn[k] = 0;
    rk[k] = fl % uint32_t(k + 1);
  }

  const bool live = f0 < n_bins;
  const uint64_t live_lanes = __ballot(live);
  const uint32_t live_bits = live ? 0xffu : 0u;  // bins past the end read as -inf (or lie in the clamped copy)
  RowLoads<H> a, b;
  hsum_load_row<H, VEC>(a, fop, n_bins, c, r_begin, fl);
  for (uint32_t r = r_begin;; r += 2) {
    hsum_load_row<H, VEC>(b, fop, n_bins, c, min(r + 1, r_last), fl);
    hsum_row<H>(st, job, a, r, fl, rk, thr, thr_key, live_lanes, live_bits, wn, wave);
    if (r + 1 > r_last) break;
    hsum_load_row<H, VEC>(a, fop, n_bins, c, min(r + 2, r_last), fl);
    hsum_row<H>(st, job, b, r + 1, fl, rk, thr, thr_key, live_lanes, live_bits, wn, wave);
    if (r + 2 > r_last) break;
  }
  // Flush what is left: one global atomic per plane for the whole workgroup.
  if (lane_id() == 0) {
#pragma unroll
    for (int k = 0; k < H; ++k) st.n[wave][k] = wn[k];
  }
  __syncthreads();
  if (tid < H) {
    uint32_t total = 0;
    for (uint32_t w = 0; w < uint32_t(kHsumThreads / 64); ++w) {
      st.base[w][tid] = total;
      total += st.n[w][tid];
    }
    const uint32_t g = total ? gatomic_add(job.counts + tid, total) : 0u;
    for (uint32_t w = 0; w < uint32_t(kHsumThreads / 64); ++w) st.base[w][tid] += g;
  }
  __syncthreads();
  const uint32_t lane = lane_id();
#pragma unroll
  for (int k = 0; k < H; ++k) {
    const uint32_t g = st.base[wave][k] + lane;
    if (lane < wn[k] && g < job.spill_cap) {
      gst(job.spill_key, size_t(k) * job.spill_cap + g, st.key[wave][k][lane]);
      gst(job.spill_pwr, size_t(k) * job.spill_cap + g, st.pwr[wave][k][lane]);
    }
  }
}

template <int H, bool VEC>
__global__ __launch_bounds__(kHsumThreads) void hsum_kernel(const HsumJob* __restrict__ jobs, uint32_t n_jobs) {
  __shared__ HsumStage<H> st;
  hsum_block<H, VEC>(st, jobs, n_jobs, blockIdx.x);
}

template <int H>
static void launch_hsum_h(const HsumJob* d_jobs, uint32_t n_jobs, uint32_t n_blocks, bool vec, hipStream_t stream) {
  if (vec)
    hipLaunchKernelGGL((hsum_kernel<H, true>), dim3(n_blocks), dim3(kHsumThreads), 0, stream, d_jobs, n_jobs);
  else
    hipLaunchKernelGGL((hsum_kernel<H, false>), dim3(n_blocks), dim3(kHsumThreads), 0, stream, d_jobs, n_jobs);
}

hipError_t launch_hsum(const HsumJob* d_jobs, uint32_t n_jobs, uint32_t n_blocks, uint32_t n_harmonics,
                       bool vec, hipStream_t stream) {
  if (n_blocks == 0) return hipSuccess;
  switch (n_harmonics) {
    case 1: launch_hsum_h<1>(d_jobs, n_jobs, n_blocks, vec, stream); break;
    case 2: launch_hsum_h<2>(d_jobs, n_jobs, n_blocks, vec, stream); break;
    case 3: launch_hsum_h<3>(d_jobs, n_jobs, n_blocks, vec, stream); break;
    case 4: launch_hsum_h<4>(d_jobs, n_jobs, n_blocks, vec, stream); break;
    case 5: launch_hsum_h<5>(d_jobs, n_jobs, n_blocks, vec, stream); break;
    case 6: launch_hsum_h<6>(d_jobs, n_jobs, n_blocks, vec, stream); break;
    case 7: launch_hsum_h<7>(d_jobs, n_jobs, n_blocks, vec, stream); break;
    case 8: launch_hsum_h<8>(d_jobs, n_jobs, n_blocks, vec, stream); break;
    case 9: launch_hsum_h<9>(d_jobs, n_jobs, n_blocks, vec, stream); break;
    case 10: launch_hsum_h<10>(d_jobs, n_jobs, n_blocks, vec, stream); break;
    case 11: launch_hsum_h<11>(d_jobs, n_jobs, n_blocks, vec, stream); break;
    case 12: launch_hsum_h<12>(d_jobs, n_jobs, n_blocks, vec, stream); break;
    case 13: launch_hsum_h<13>(d_jobs, n_jobs, n_blocks, vec, stream); break;
    case 14: launch_hsum_h<14>(d_jobs, n_jobs, n_blocks, vec, stream); break;
    case 15: launch_hsum_h<15>(d_jobs, n_jobs, n_blocks, vec, stream); break;
    case 16: launch_hsum_h<16>(d_jobs, n_jobs, n_blocks, vec, stream); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Fused launch: FT convolution of one group with 1024-point tiles and the
// harmonic summing of the group before it
// ---------------------------------------------------------------------------

// Workgroup v is either an FT-convolution work item or a summing block; the
// two kinds are interleaved evenly over the grid (v is an FT item iff
// floor((v+1) A / (A+B)) > floor(v A / (A+B)) for A items and B blocks), so
// the hardware's in-order dispatch keeps both in flight together: the
// VALU-bound transforms and the HBM-bound summing share every CU instead of
// taking turns.  The summing reads a FOP that an earlier launch wrote, so no
// workgroup waits on another.  LDS is one buffer viewed as either kind's.
template <int H>
union FusedLds {
  Ftc1kLds ftc;
  HsumStage<H> hsum;
};

template <int H, bool VEC>
__global__ __launch_bounds__(64 * kFtc1kWaves) __attribute__((amdgpu_waves_per_eu(3, 3))) void fused1k_kernel(
    const FtcJob* __restrict__ ftc_jobs, uint32_t n_ftc_jobs, uint32_t n_items, const float2* __restrict__ tw1,
    const float2* __restrict__ tw2, const HsumJob* __restrict__ hsum_jobs, uint32_t n_hsum_jobs,
    uint32_t n_blocks) {
  static_assert(kHsumThreads == 64 * kFtc1kWaves, "both kinds of work use the same workgroup size");
  __shared__ FusedLds<H> lds;
  const uint32_t v = blockIdx.x;
  const uint64_t total = uint64_t(n_items) + n_blocks;
  const uint32_t a0 = uint32_t(uint64_t(v) * n_items / total);
  const uint32_t a1 = uint32_t(uint64_t(v + 1) * n_items / total);
  if (a1 > a0) {
    WaveFFT1k<true> fft;
    fft.load_twiddles(tw1, tw2, lds.ftc.w2s, int(threadIdx.x) & 63, (threadIdx.x >> 6) == 0);
    ftc1k_item(lds.ftc, fft, ftc_jobs, n_ftc_jobs, a0);
  } else {
    hsum_block<H, VEC>(lds.hsum, hsum_jobs, n_hsum_jobs, v - a0);
  }
}

bool fused_supported(uint32_t tile, uint32_t n_harmonics, bool vec) {
  return tile == 1024 && (n_harmonics == 8 || (n_harmonics == 16 && vec));
}

hipError_t launch_fused(const FtcJob* ftc_jobs, uint32_t n_ftc_jobs, uint32_t n_items, const float2* tw1,
                        const float2* tw2, const HsumJob* hsum_jobs, uint32_t n_hsum_jobs, uint32_t n_blocks,
                        uint32_t n_harmonics, bool vec, hipStream_t stream) {
  const uint32_t grid = n_items + n_blocks;
  if (grid == 0) return hipSuccess;
  const dim3 block(64 * kFtc1kWaves);
#define FDAS_FUSED(HH, VV)                                                                                          \
  hipLaunchKernelGGL((fused1k_kernel<HH, VV>), dim3(grid), block, 0, stream, ftc_jobs, n_ftc_jobs, n_items, tw1, tw2, \
                     hsum_jobs, n_hsum_jobs, n_blocks)
  if (n_harmonics == 8 && vec) FDAS_FUSED(8, true);
  else if (n_harmonics == 8) FDAS_FUSED(8, false);
  else if (n_harmonics == 16 && vec) FDAS_FUSED(16, true);
  else return hipErrorInvalidValue;
#undef FDAS_FUSED
  return hipGetLastError();
}

__global__ __launch_bounds__(256) void hplane_kernel(const float* __restrict__ fop, uint32_t n_bins,
                                                     uint32_t n_templates, uint32_t row_begin, uint32_t k_max,
                                                     float* __restrict__ out) {
  const uint32_t f = blockIdx.x * 256u + threadIdx.x;
  const uint32_t r = row_begin + blockIdx.y;
  if (f >= n_bins) return;
  const int c = int(n_templates / 2);
  const int t = int(r) - c;
  float hp = fop[size_t(r) * n_bins + f];
  for (uint32_t k = 2; k <= k_max; ++k) {
    const int rk = c + t / int(k);
    hp += fop[size_t(rk) * n_bins + f / k];
  }
  out[size_t(r) * n_bins + f] = hp;
}

hipError_t launch_hplane(const float* fop, uint32_t n_bins, uint32_t n_templates, uint32_t row_begin, uint32_t n_rows,
                         uint32_t k, float* out, hipStream_t stream) {
  if (n_rows == 0 || n_bins == 0) return hipSuccess;
  dim3 grid((n_bins + 255) / 256, n_rows);
  hipLaunchKernelGGL(hplane_kernel, grid, dim3(256), 0, stream, fop, n_bins, n_templates, row_begin, k, out);
  return hipGetLastError();
}

// Largest HP_k, k = 1..n_harmonics, of each block of 256 bins x kHmaxRows rows of
// rows [row_begin, row_begin + n_rows): out[k-1][block], block = y * gridDim.x + x.
// HP_k is summed in FP32 in k order, as in hsum_kernel, so the values are the
// ones the summing kernel thresholds.
constexpr uint32_t kHmaxRows = 8;

__global__ __launch_bounds__(256) void hmax_kernel(const float* __restrict__ fop, uint32_t n_bins,
                                                   uint32_t n_templates, uint32_t row_begin, uint32_t n_rows,
                                                   uint32_t n_harmonics, float* __restrict__ out) {
  __shared__ float red[kMaxHarmonics][256 / 64];
  const uint32_t f = blockIdx.x * 256u + threadIdx.x;
  const uint32_t r0 = row_begin + blockIdx.y * kHmaxRows;
  const uint32_t r1 = min(r0 + kHmaxRows, row_begin + n_rows);
  const int c = int(n_templates / 2);
  float mx[kMaxHarmonics];
  for (uint32_t k = 0; k < kMaxHarmonics; ++k) mx[k] = -INFINITY;
  if (f < n_bins) {
    for (uint32_t r = r0; r < r1; ++r) {
      const int t = int(r) - c;
      float hp = gld(fop, size_t(r) * n_bins + f);
      mx[0] = fmaxf(mx[0], hp);
      for (uint32_t k = 2; k <= n_harmonics; ++k) {
        hp += gld(fop, size_t(c + t / int(k)) * n_bins + f / k);
        mx[k - 1] = fmaxf(mx[k - 1], hp);
      }
    }
  }
  const uint32_t lane = lane_id(), wave = threadIdx.x >> 6;
  for (uint32_t k = 0; k < n_harmonics; ++k) {
    float v = mx[k];
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
    if (lane == 0) red[k][wave] = v;
  }
  __syncthreads();
  if (threadIdx.x < n_harmonics) {
    const uint32_t k = threadIdx.x;
    const float v = fmaxf(fmaxf(red[k][0], red[k][1]), fmaxf(red[k][2], red[k][3]));
    gst(out, size_t(k) * (gridDim.x * gridDim.y) + blockIdx.y * gridDim.x + blockIdx.x, v);
  }
}

uint32_t hmax_blocks(uint32_t n_bins, uint32_t n_rows) {
  return (n_bins + 255) / 256 * ((n_rows + kHmaxRows - 1) / kHmaxRows);
}

hipError_t launch_hmax(const float* fop, uint32_t n_bins, uint32_t n_templates, uint32_t row_begin, uint32_t n_rows,
                       uint32_t n_harmonics, float* out, hipStream_t stream) {
  if (n_rows == 0 || n_bins == 0) return hipSuccess;
  dim3 grid((n_bins + 255) / 256, (n_rows + kHmaxRows - 1) / kHmaxRows);
  hipLaunchKernelGGL(hmax_kernel, grid, dim3(256), 0, stream, fop, n_bins, n_templates, row_begin, n_rows,
                     n_harmonics, out);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Selection: keep the `cap` strongest detections of a plane (ties: smaller
// (row, bin) key first), returned sorted by key.
// ---------------------------------------------------------------------------

// Exclusive prefix sum over a workgroup of THREADS threads; *total gets the sum.
template <int THREADS>
__device__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* s_waves, uint32_t* total) {
  constexpr int n_waves = THREADS / 64;
  const int tid = threadIdx.x;
  const int lane = int(lane_id());
  const int wave = tid >> 6;
  uint32_t x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t y = __shfl_up(x, o);
    if (lane >= o) x += y;
  }
  if (lane == 63) s_waves[wave] = x;
  __syncthreads();
  if (tid < 64) {
    const uint32_t w = tid < n_waves ? s_waves[tid] : 0u;
    uint32_t ws = w;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t y = __shfl_up(ws, o);
      if (lane >= o) ws += y;
    }
    if (tid < n_waves) s_waves[tid] = ws - w;
    if (tid == n_waves - 1) s_waves[32] = ws;
  }
  __syncthreads();
  const uint32_t res = s_waves[wave] + x - v;
  *total = s_waves[32];
  __syncthreads();
  return res;
}

// Cut bin b*: the highest bin whose suffix count sum_{b >= b*} hist[b] reaches
// `cap`.  s_out[0] = b*, s_out[1] = entries strictly above b*.  The histogram
// counts every spilled entry and there are more than `cap` of them, so a cut
// exists; the defaults (bin 0, nothing above) keep b* inside the histogram
// regardless.
template <int THREADS>
__device__ void find_cut(const uint32_t* __restrict__ hist, uint32_t cap, uint32_t* s_waves, uint32_t* s_out) {
  constexpr int PER = int(kSelectBins) / THREADS;
  const int tid = threadIdx.x;
  if (tid == 0) {
    s_out[0] = 0;
    s_out[1] = 0;
  }
  uint32_t h[PER];
  uint32_t local = 0;
#pragma unroll
  for (int j = 0; j < PER; ++j) {
    h[j] = gld(hist, kSelectBins - 1 - uint32_t(tid * PER + j));
    local += h[j];
  }
  uint32_t total;
  const uint32_t excl = block_exclusive_scan<THREADS>(local, s_waves, &total);
  uint32_t bin, above;
  if (cut_in_share<PER>(h, excl, cap, uint32_t(tid * PER), &bin, &above)) {
    s_out[0] = bin;
    s_out[1] = above;
  }
  __syncthreads();
}

__global__ __launch_bounds__(256) void cand_hist_kernel(const SelectJob* __restrict__ jobs) {
  const SelectJob& job = jobs[blockIdx.y];
  const uint32_t n = min(gld(job.count, 0), job.spill_cap);
  if (n <= job.cap) return;
  __shared__ uint32_t h[kSelectBins];
  const int tid = threadIdx.x;
  for (uint32_t i = tid; i < kSelectBins; i += 256) h[i] = 0;
  __syncthreads();
  const uint32_t tb = histogram_base(__float_as_uint(job.thr));
  const uint32_t per = (n + kSelectSlices - 1) / kSelectSlices;
  const uint32_t beg = blockIdx.x * per, end = min(beg + per, n);
  const float* __restrict__ pw = job.spill_pwr;
  for (uint32_t e0 = beg; e0 < end; e0 += 4 * 256) {
    uint32_t v[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const uint32_t e = e0 + uint32_t(u * 256 + tid);
      v[u] = e < end ? __float_as_uint(gld(pw, e)) : 0xffffffffu;  // all ones: past the slice
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
      if (v[u] != 0xffffffffu) atomicAdd(&h[power_bin(v[u], tb)], 1u);
  }
  __syncthreads();
  for (uint32_t i = tid; i < kSelectBins; i += 256)
    if (h[i]) gatomic_add(job.hist + i, h[i]);
}

__global__ __launch_bounds__(256) void cand_compact_kernel(const SelectJob* __restrict__ jobs) {
  const SelectJob& job = jobs[blockIdx.y];
  const uint32_t n = min(gld(job.count, 0), job.spill_cap);
  if (n <= job.cap) return;
  __shared__ uint32_t waves[40];
  __shared__ uint32_t cut[2];
  find_cut<256>(job.hist, job.cap, waves, cut);
  const uint32_t bstar = cut[0];
  const uint32_t tb = histogram_base(__float_as_uint(job.thr));
  const int tid = threadIdx.x;
  const uint32_t per = (n + kSelectSlices - 1) / kSelectSlices;
  const uint32_t beg = blockIdx.x * per, end = min(beg + per, n);
  for (uint32_t e0 = beg; e0 < end; e0 += 256) {
    const uint32_t e = e0 + uint32_t(tid);
    float v = -1.f;
    uint32_t key = 0;
    if (e < end) {
      v = gld(job.spill_pwr, e);
      key = gld(job.spill_key, e);
    }
    const bool sel = e < end && power_bin(__float_as_uint(v), tb) >= bstar;
    const uint64_t mask = __ballot(sel);
    if (mask == 0) continue;
    const uint32_t lane = lane_id();
    const uint32_t leader = uint32_t(__builtin_ctzll(mask));
    uint32_t pos = 0;
    if (lane == leader) pos = gatomic_add(job.compact_n, uint32_t(__builtin_popcountll(mask)));
    pos = __shfl(pos, int(leader)) + lanes_below(mask);
    if (sel && pos < job.compact_cap) {
      gst(job.compact_key, pos, key);
      gst(job.compact_pwr, pos, v);
    }
  }
}

struct SelectShared {
  uint32_t key[kSelectMaxCap];
  float pwr[kSelectMaxCap];
  uint64_t pool[kSelectPool];
  uint32_t hist[2048];
  uint32_t waves[40];
  uint32_t misc[8];
};

// One radix-select digit pass: among values v with (v & mask) == prefix, find
// the digit (bits [shift, shift+width)) holding the `rank`-th element counted
// from the top (descending) or bottom (ascending).  Updates prefix/mask/rank.
template <bool DESCENDING, typename Get>
__device__ void radix_pass(SelectShared& sh, uint32_t n, Get get, uint32_t shift, uint32_t width, uint32_t& prefix,
                           uint32_t& mask, uint32_t& rank) {
  const int tid = threadIdx.x;
  const uint32_t nb = 1u << width;
  for (uint32_t i = tid; i < 2048; i += kSelectThreads) sh.hist[i] = 0;
  __syncthreads();
  for (uint32_t e = tid; e < n; e += kSelectThreads) {
    const uint32_t v = get(e);
    if ((v & mask) == prefix) atomicAdd(&sh.hist[(v >> shift) & (nb - 1)], 1u);
  }
  __syncthreads();
  // Thread t owns ordered positions u = 2t, 2t+1 (u counts from the selected end).
  uint32_t h[2];
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    const uint32_t u = uint32_t(2 * tid + j);
    const uint32_t d = DESCENDING ? (nb - 1 - u) : u;
    h[j] = u < nb ? sh.hist[d] : 0u;
  }
  uint32_t total;
  const uint32_t excl = block_exclusive_scan<kSelectThreads>(h[0] + h[1], sh.waves, &total);
  if (excl < rank && rank <= excl + h[0] + h[1]) {
    uint32_t u = uint32_t(2 * tid);
    uint32_t r = rank - excl;
    if (r > h[0]) {
      r -= h[0];
      u += 1;
    }
    const uint32_t d = DESCENDING ? (nb - 1 - u) : u;
    sh.misc[0] = d;
    sh.misc[1] = r;
    sh.misc[2] = sh.hist[d];
  }
  __syncthreads();
  prefix |= sh.misc[0] << shift;
  mask |= (nb - 1) << shift;
  rank = sh.misc[1];
  __syncthreads();
}

__device__ void bitonic_sort_by_key(SelectShared& sh, uint32_t n) {
  const int tid = threadIdx.x;
  uint32_t p2 = 1;
  while (p2 < n) p2 <<= 1;
  for (uint32_t i = n + tid; i < p2; i += kSelectThreads) {
    sh.key[i] = 0xffffffffu;
    sh.pwr[i] = 0.f;
  }
  __syncthreads();
  for (uint32_t k = 2; k <= p2; k <<= 1) {
    for (uint32_t j = k >> 1; j > 0; j >>= 1) {
      for (uint32_t i = tid; i < p2; i += kSelectThreads) {
        const uint32_t ixj = i ^ j;
        if (ixj > i) {
          const uint32_t a = sh.key[i], b = sh.key[ixj];
          const bool up = (i & k) == 0;
          if ((a > b) == up) {
            sh.key[i] = b;
            sh.key[ixj] = a;
            const float pa = sh.pwr[i];
            sh.pwr[i] = sh.pwr[ixj];
            sh.pwr[ixj] = pa;
          }
        }
      }
      __syncthreads();
    }
  }
}

__device__ void bitonic_sort_u64(uint64_t* a, uint32_t n) {
  const int tid = threadIdx.x;
  uint32_t p2 = 1;
  while (p2 < n) p2 <<= 1;
  for (uint32_t i = n + tid; i < p2; i += kSelectThreads) a[i] = ~0ull;
  __syncthreads();
  for (uint32_t k = 2; k <= p2; k <<= 1) {
    for (uint32_t j = k >> 1; j > 0; j >>= 1) {
      for (uint32_t i = tid; i < p2; i += kSelectThreads) {
        const uint32_t ixj = i ^ j;
        if (ixj > i) {
          const uint64_t x = a[i], y = a[ixj];
          if ((x > y) == ((i & k) == 0)) {
            a[i] = y;
            a[ixj] = x;
          }
        }
      }
      __syncthreads();
    }
  }
}

// Exact top-`cap` of the whole spill by radix select on (power desc, key asc).
// Used when the boundary bin holds more entries than the LDS pool.
__device__ uint32_t radix_select_spill(SelectShared& sh, const SelectJob& job, uint32_t n) {
  const int tid = threadIdx.x;
  const uint32_t cap = job.cap;
  const uint32_t* __restrict__ skey = job.spill_key;
  const float* __restrict__ spwr = job.spill_pwr;
  auto pkey = [&](uint32_t e) { return order_key(__float_as_uint(gld(spwr, e))); };
  uint32_t prefix = 0, mask = 0, rank = cap;
  radix_pass<true>(sh, n, pkey, 21, 11, prefix, mask, rank);
  radix_pass<true>(sh, n, pkey, 10, 11, prefix, mask, rank);
  radix_pass<true>(sh, n, pkey, 0, 10, prefix, mask, rank);
  const uint32_t vkey = prefix;  // order key of the cap-th strongest power V
  const uint32_t n_equal = sh.misc[2];
  // `rank` entries with power == V are kept: the smallest keys among them.
  uint32_t kmax = 0xffffffffu;
  if (n_equal > rank) {
    auto kbits = [&](uint32_t e) { return pkey(e) == vkey ? gld(skey, e) : 0xffffffffu; };
    uint32_t kp = 0, km = 0, kr = rank;
    radix_pass<false>(sh, n, kbits, 21, 11, kp, km, kr);
    radix_pass<false>(sh, n, kbits, 10, 11, kp, km, kr);
    radix_pass<false>(sh, n, kbits, 0, 10, kp, km, kr);
    kmax = kp;
  }
  if (tid == 0) sh.misc[3] = 0;
  __syncthreads();
  for (uint32_t e = tid; e < n; e += kSelectThreads) {
    const float pv = gld(spwr, e);
    const uint32_t pk = order_key(__float_as_uint(pv));
    const uint32_t key = gld(skey, e);
    if (pk > vkey || (pk == vkey && key <= kmax)) {
      const uint32_t pos = atomicAdd(&sh.misc[3], 1u);
      if (pos < cap) {
        sh.key[pos] = key;
        sh.pwr[pos] = pv;
      }
    }
  }
  __syncthreads();
  return min(sh.misc[3], cap);
}

__global__ __launch_bounds__(kSelectThreads) void select_final_kernel(const SelectJob* __restrict__ jobs) {
  __shared__ SelectShared sh;
  const SelectJob& job = jobs[blockIdx.x];
  const int tid = threadIdx.x;
  const uint32_t n = min(gld(job.count, 0), job.spill_cap);
  const uint32_t cap = job.cap;

  uint32_t kept;
  if (n <= cap) {
    for (uint32_t e = tid; e < n; e += kSelectThreads) {
      sh.key[e] = gld(job.spill_key, e);
      sh.pwr[e] = gld(job.spill_pwr, e);
    }
    kept = n;
  } else {
    find_cut<kSelectThreads>(job.hist, cap, sh.waves, sh.misc);
    const uint32_t bstar = sh.misc[0], above = sh.misc[1];
    const uint32_t m = gld(job.compact_n, 0);
    const uint32_t pool_n = gld(job.hist, bstar);
    __syncthreads();
    if (m <= job.compact_cap && pool_n <= kSelectPool && above < cap) {
      const uint32_t tb = histogram_base(__float_as_uint(job.thr));
      if (tid == 0) {
        sh.misc[3] = 0;
        sh.misc[4] = 0;
      }
      __syncthreads();
      for (uint32_t e = tid; e < m; e += kSelectThreads) {
        const float v = gld(job.compact_pwr, e);
        const uint32_t key = gld(job.compact_key, e);
        const uint32_t pb = __float_as_uint(v);
        const uint32_t bin = power_bin(pb, tb);
        if (bin > bstar) {
          const uint32_t pos = atomicAdd(&sh.misc[3], 1u);
          if (pos < cap) {
            sh.key[pos] = key;
            sh.pwr[pos] = v;
          }
        } else if (bin == bstar) {
          const uint32_t q = atomicAdd(&sh.misc[4], 1u);
          if (q < kSelectPool) sh.pool[q] = (uint64_t(~order_key(pb)) << 32) | key;  // strongest first
        }
      }
      __syncthreads();
      const uint32_t n_pool = min(sh.misc[4], kSelectPool);
      bitonic_sort_u64(sh.pool, n_pool);
      const uint32_t take = cap - above;
      for (uint32_t i = tid; i < take && i < n_pool; i += kSelectThreads) {
        const uint64_t v = sh.pool[i];
        sh.key[above + i] = uint32_t(v);
        sh.pwr[above + i] = __uint_as_float(order_key_inv(~uint32_t(v >> 32)));
      }
      __syncthreads();
      kept = cap;
    } else {
      kept = radix_select_spill(sh, job, n);
    }
  }
  __syncthreads();
  bitonic_sort_by_key(sh, kept);
  for (uint32_t e = tid; e < kept; e += kSelectThreads) {
    gst(job.out_key, e, sh.key[e]);
    gst(job.out_pwr, e, sh.pwr[e]);
  }
  if (tid == 0) gst(job.out_n, 0, kept);
}

hipError_t launch_select(const SelectJob* d_jobs, uint32_t n_jobs, hipStream_t stream) {
  if (n_jobs == 0) return hipSuccess;
  hipLaunchKernelGGL(cand_hist_kernel, dim3(kSelectSlices, n_jobs), dim3(256), 0, stream, d_jobs);
  hipLaunchKernelGGL(cand_compact_kernel, dim3(kSelectSlices, n_jobs), dim3(256), 0, stream, d_jobs);
  hipLaunchKernelGGL(select_final_kernel, dim3(n_jobs), dim3(kSelectThreads), 0, stream, d_jobs);
  return hipGetLastError();
}

__global__ void zero_u32_kernel(uint32_t* __restrict__ p, uint32_t n) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0u;
}

hipError_t launch_zero_u32(uint32_t* p, uint32_t n, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  const uint32_t blocks = min((n + 255u) / 256u, 1024u);
  hipLaunchKernelGGL(zero_u32_kernel, dim3(blocks), dim3(256), 0, stream, p, n);
  return hipGetLastError();
}

}  // namespace fdas
