// Device front end: dedispersed time series -> the spectrum the search takes,
// X[f] = DFT_Ns(s)[f] / Ns for f < N, Ns = 2N (ft_conv.py:123-146), in FP32 on
// the GPU (the reference computes it in FP32 too: scipy.fft.fft of float32
// samples returns complex64).
//
// The 2N real samples are read as N complex points z[n] = s[2n] + i s[2n+1]
// (the float32 array reinterpreted, no copy).  Z = DFT_N(z) uses the four-step
// factorisation N = N1 * N2 with explicit transposes so every pass reads and
// writes contiguous rows:
//   z [N1][N2] -T-> [N2][N1] -row DFT_N1, x W_N^(n2 k1)-> -T-> [N1][N2]
//   -row DFT_N2-> [N1][N2] -T-> Z [N2][N1] (natural order)
// Rows of 1024, 2048 or 4096 points use the tile transform of the FT
// convolution; other lengths (N2 < 1024, or N <= 4096) a radix-2 LDS kernel.
// Finally X[f] = (E[f] + W_Ns^f O[f]) / Ns with E = (Z[f] + conj Z[N-f]) / 2 and
// O = (Z[f] - conj Z[N-f]) / (2i), the even/odd-sample split of a real DFT.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <string>
#include <vector>

#include "fdas_host.h"
#include "fdas_plan.h"
#include "fft_tile.h"

namespace fdas {
namespace {

constexpr int kTr = 32;  // transpose tile

// out[c][r] = in[r][c] for an R x C matrix of complex64.
__global__ __launch_bounds__(kTr * 8) void transpose_c64_kernel(const float2* __restrict__ in, float2* __restrict__ out,
                                                                uint32_t R, uint32_t C) {
  __shared__ float2 t[kTr][kTr + 1];
  const uint32_t c0 = blockIdx.x * kTr, r0 = blockIdx.y * kTr;
  const uint32_t tx = threadIdx.x, ty = threadIdx.y;
  for (uint32_t j = ty; j < kTr; j += 8) {
    const uint32_t r = r0 + j, c = c0 + tx;
    if (r < R && c < C) t[j][tx] = gld2(in, size_t(r) * C + c);
  }
  __syncthreads();
  for (uint32_t j = ty; j < kTr; j += 8) {
    const uint32_t c = c0 + j, r = r0 + tx;
    if (r < R && c < C) {
      const float2 v = t[tx][j];
      ((FDAS_GLOBAL fdas_vf2*)out)[size_t(c) * R + r] = fdas_vf2{v.x, v.y};
    }
  }
}

// e^(-2 pi i m / 2^log2n) for 0 <= m < 2^log2n (<= 2^23): the argument -2m/2^log2n
// is exact in FP32, so the only rounding is sincospi's own.
__device__ __forceinline__ float2 root_of_unity(uint32_t m, int log2n) {
  float s, c;
  sincospif(-ldexpf(float(m), 1 - log2n), &s, &c);
  return make_float2(c, s);
}

// One S-point forward DFT per row (S = 1024, 2048, 4096), rows contiguous; with
// log2n > 0 output bin k of row r is multiplied by W_N^(r k), N = 2^log2n.
template <int S>
__global__ __launch_bounds__(S / 16) void rowfft_kernel(const float2* __restrict__ in, float2* __restrict__ out,
                                                        uint32_t n_rows, const float2* __restrict__ tw1,
                                                        const float2* __restrict__ tw2, int log2n) {
  using F = TileFFT<S>;
  constexpr int NT = F::NT;
  __shared__ float2 la[F::LDS_ELEMS];
  const int tid = threadIdx.x;
  F fft;
  fft.load_twiddles(tw1, tw2, tid);
  for (uint32_t row = blockIdx.x; row < n_rows; row += gridDim.x) {
    __syncthreads();
    const float2* __restrict__ x = in + size_t(row) * S;
    float2 v[16];
#pragma unroll
    for (int a = 0; a < 16; ++a) v[a] = gld2(x, size_t(NT * a + tid));
    fft.forward(v, la, tid);
    float2* __restrict__ o = out + size_t(row) * S;
#pragma unroll
    for (int s = 0; s < 16; ++s) {
      const uint32_t k = uint32_t(F::slot_bin(tid, s));
      float2 y = v[s];
      if (log2n > 0) y = cmul(y, root_of_unity((row * k) & ((1u << log2n) - 1u), log2n));
      ((FDAS_GLOBAL fdas_vf2*)o)[k] = fdas_vf2{y.x, y.y};
    }
  }
}

constexpr int kLdsFftMax = 4096;
constexpr int kLdsFftThreads = 256;

__device__ __forceinline__ uint32_t bit_reverse(uint32_t i, int bits) { return __brev(i) >> (32 - bits); }

// One L-point forward DFT per row (L = 2^log2l <= 4096), radix-2 in LDS; the
// optional row twiddle as in rowfft_kernel.
__global__ __launch_bounds__(kLdsFftThreads) void ldsfft_kernel(const float2* __restrict__ in, float2* __restrict__ out,
                                                                uint32_t n_rows, int log2l, int log2n) {
  __shared__ float2 buf[kLdsFftMax];
  const uint32_t L = 1u << log2l, tid = threadIdx.x;
  for (uint32_t row = blockIdx.x; row < n_rows; row += gridDim.x) {
    __syncthreads();
    for (uint32_t i = tid; i < L; i += kLdsFftThreads)
      buf[log2l ? bit_reverse(i, log2l) : 0] = gld2(in, size_t(row) * L + i);
    __syncthreads();
    for (int s = 1; s <= log2l; ++s) {
      const uint32_t half = 1u << (s - 1);
      for (uint32_t b = tid; b < L / 2; b += kLdsFftThreads) {
        const uint32_t pos = b & (half - 1u);
        const uint32_t i0 = ((b >> (s - 1)) << s) + pos, i1 = i0 + half;
        const float2 w = root_of_unity(pos, s);
        const float2 t = cmul(w, buf[i1]);
        const float2 u = buf[i0];
        buf[i0] = cadd(u, t);
        buf[i1] = csub(u, t);
      }
      __syncthreads();
    }
    for (uint32_t k = tid; k < L; k += kLdsFftThreads) {
      float2 y = buf[k];
      if (log2n > 0) y = cmul(y, root_of_unity((row * k) & ((1u << log2n) - 1u), log2n));
      ((FDAS_GLOBAL fdas_vf2*)out)[size_t(row) * L + k] = fdas_vf2{y.x, y.y};
    }
  }
}

// X[f] = (E + W_Ns^f O) / Ns from the N-point transform Z of the packed series.
__global__ __launch_bounds__(256) void real_spectrum_kernel(const float2* __restrict__ Z, float2* __restrict__ X,
                                                            uint32_t n, int log2n, float inv_ns) {
  for (uint32_t f = blockIdx.x * blockDim.x + threadIdx.x; f < n; f += gridDim.x * blockDim.x) {
    const float2 a = gld2(Z, f);
    const float2 b = gld2(Z, (n - f) & (n - 1u));  // Z[N - f], Z[N] = Z[0]
    // E = (a + conj b) / 2, O = (a - conj b) / (2i)
    const float2 e = make_float2(0.5f * (a.x + b.x), 0.5f * (a.y - b.y));
    const float2 o = make_float2(0.5f * (a.y + b.y), -0.5f * (a.x - b.x));
    const float2 w = root_of_unity(f, log2n + 1);  // W_Ns^f, Ns = 2N
    const float2 x = cadd(e, cmul(w, o));
    ((FDAS_GLOBAL fdas_vf2*)X)[f] = fdas_vf2{x.x * inv_ns, x.y * inv_ns};
  }
}

int ilog2(uint32_t v) {
  int l = 0;
  while ((1u << l) < v) ++l;
  return l;
}

}  // namespace
}  // namespace fdas

using namespace fdas;

struct fdas_frontend {
  uint32_t n = 0, n1 = 0, n2 = 0;
  int device = 0;
  std::string last_error;
  hipStream_t stream = nullptr;
  float2* buf[2] = {nullptr, nullptr};
  float2* series = nullptr;    // device copy of a host series
  float2* spectrum = nullptr;  // device result for a host destination
  float2* tw1[3] = {nullptr, nullptr, nullptr};  // per tile size 1024 / 2048 / 4096
  float2* tw2[3] = {nullptr, nullptr, nullptr};
};

namespace {

fdas_status fe_fail(fdas_frontend* fe, fdas_status st, const char* fmt, ...) {
  char b[512];
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(b, sizeof(b), fmt, ap);
  va_end(ap);
  if (fe) fe->last_error = std::string("[ERROR] ") + b;
  return st;
}

#define FEHIP(fe, expr)                                                                                          \
  do {                                                                                                           \
    hipError_t e_ = (expr);                                                                                      \
    if (e_ != hipSuccess)                                                                                        \
      return fe_fail(fe, FDAS_EDEVICE, "HIP call failed with status %d (%s): %s", int(e_), hipGetErrorString(e_), \
                     #expr);                                                                                     \
  } while (0)

int tile_index(uint32_t s) { return s == 1024 ? 0 : s == 2048 ? 1 : s == 4096 ? 2 : -1; }

// Row DFTs of length L over n_rows contiguous rows, optional W_N^(row k) twiddle.
fdas_status row_dfts(fdas_frontend* fe, const float2* in, float2* out, uint32_t L, uint32_t n_rows, int log2n) {
  const uint32_t grid = std::min<uint32_t>(n_rows, 65535);
  const int ti = tile_index(L);
  switch (ti) {
    case 0:
      hipLaunchKernelGGL(rowfft_kernel<1024>, dim3(grid), dim3(64), 0, fe->stream, in, out, n_rows, fe->tw1[0],
                         fe->tw2[0], log2n);
      break;
    case 1:
      hipLaunchKernelGGL(rowfft_kernel<2048>, dim3(grid), dim3(128), 0, fe->stream, in, out, n_rows, fe->tw1[1],
                         fe->tw2[1], log2n);
      break;
    case 2:
      hipLaunchKernelGGL(rowfft_kernel<4096>, dim3(grid), dim3(256), 0, fe->stream, in, out, n_rows, fe->tw1[2],
                         fe->tw2[2], log2n);
      break;
    default:
      hipLaunchKernelGGL(ldsfft_kernel, dim3(grid), dim3(kLdsFftThreads), 0, fe->stream, in, out, n_rows, ilog2(L),
                         log2n);
  }
  FEHIP(fe, hipGetLastError());
  return FDAS_OK;
}

fdas_status transpose(fdas_frontend* fe, const float2* in, float2* out, uint32_t R, uint32_t C) {
  const dim3 grid((C + kTr - 1) / kTr, (R + kTr - 1) / kTr);
  hipLaunchKernelGGL(transpose_c64_kernel, grid, dim3(kTr, 8), 0, fe->stream, in, out, R, C);
  FEHIP(fe, hipGetLastError());
  return FDAS_OK;
}

fdas_status fe_init(fdas_frontend* fe) {
  FEHIP(fe, hipSetDevice(fe->device));
  FEHIP(fe, hipStreamCreateWithFlags(&fe->stream, hipStreamNonBlocking));
  for (int i = 0; i < 2; ++i) FEHIP(fe, hipMalloc(&fe->buf[i], sizeof(float2) * fe->n));
  for (uint32_t s : {1024u, 2048u, 4096u}) {
    if (fe->n1 != s && fe->n2 != s) continue;
    std::vector<float2> t1, t2;
    make_twiddles(s, t1, t2);
    const int i = tile_index(s);
    FEHIP(fe, hipMalloc(&fe->tw1[i], sizeof(float2) * t1.size()));
    FEHIP(fe, hipMalloc(&fe->tw2[i], sizeof(float2) * t2.size()));
    FEHIP(fe, hipMemcpy(fe->tw1[i], t1.data(), sizeof(float2) * t1.size(), hipMemcpyHostToDevice));
    FEHIP(fe, hipMemcpy(fe->tw2[i], t2.data(), sizeof(float2) * t2.size(), hipMemcpyHostToDevice));
  }
  return FDAS_OK;
}

}  // namespace

extern "C" {

fdas_status fdas_frontend_create(uint32_t n_bins, int device, fdas_frontend** out) {
  if (!out) return FDAS_EINVAL;
  *out = nullptr;
  if (n_bins < 2 || n_bins > (1u << 22) || (n_bins & (n_bins - 1u))) return FDAS_EINVAL;
  int n_dev = 0;
  if (hipGetDeviceCount(&n_dev) != hipSuccess || device < 0 || device >= n_dev) {
    (void)hipGetLastError();
    return FDAS_EDEVICE;
  }
  auto* fe = new fdas_frontend();
  fe->n = n_bins;
  fe->device = device;
  if (n_bins <= 4096) {
    fe->n1 = n_bins;
    fe->n2 = 1;
  } else {
    fe->n1 = n_bins >= (1u << 21) ? 2048u : 1024u;
    fe->n2 = n_bins / fe->n1;
  }
  if (fe_init(fe) != FDAS_OK) {
    fdas_frontend_destroy(fe);
    return FDAS_EDEVICE;
  }
  *out = fe;
  return FDAS_OK;
}

void fdas_frontend_destroy(fdas_frontend* fe) {
  if (!fe) return;
  (void)hipSetDevice(fe->device);
  if (fe->stream) (void)hipStreamSynchronize(fe->stream);
  for (float2* p : {fe->buf[0], fe->buf[1], fe->series, fe->spectrum, fe->tw1[0], fe->tw1[1], fe->tw1[2], fe->tw2[0],
                    fe->tw2[1], fe->tw2[2]})
    (void)hipFree(p);
  if (fe->stream) (void)hipStreamDestroy(fe->stream);
  delete fe;
}

const char* fdas_frontend_last_error(const fdas_frontend* fe) { return fe ? fe->last_error.c_str() : "null front end"; }

fdas_status fdas_frontend_run(fdas_frontend* fe, const float* series, uint32_t flags, void* spectrum_c64) {
  if (!fe) return FDAS_EINVAL;
  if (!series || !spectrum_c64) return fe_fail(fe, FDAS_EINVAL, "null series or spectrum");
  FEHIP(fe, hipSetDevice(fe->device));
  const uint32_t N = fe->n, N1 = fe->n1, N2 = fe->n2;
  const int log2n = ilog2(N);
  const float2* z = reinterpret_cast<const float2*>(series);
  if (!(flags & FDAS_FE_DEVICE_INPUT)) {
    if (!fe->series) FEHIP(fe, hipMalloc(&fe->series, sizeof(float2) * N));
    FEHIP(fe, hipMemcpyAsync(fe->series, series, sizeof(float2) * N, hipMemcpyHostToDevice, fe->stream));
    z = fe->series;
  }
  float2* X = static_cast<float2*>(spectrum_c64);
  if (!(flags & FDAS_FE_DEVICE_OUTPUT)) {
    if (!fe->spectrum) FEHIP(fe, hipMalloc(&fe->spectrum, sizeof(float2) * N));
    X = fe->spectrum;
  }
  float2 *b0 = fe->buf[0], *b1 = fe->buf[1];
  fdas_status st;
  const float2* Z = nullptr;
  if (N2 == 1) {
    if ((st = row_dfts(fe, z, b0, N, 1, 0)) != FDAS_OK) return st;
    Z = b0;
  } else {
    if ((st = transpose(fe, z, b0, N1, N2)) != FDAS_OK) return st;        // [N2][N1]
    if ((st = row_dfts(fe, b0, b1, N1, N2, log2n)) != FDAS_OK) return st;  // DFT over n1, x W_N^(n2 k1)
    if ((st = transpose(fe, b1, b0, N2, N1)) != FDAS_OK) return st;       // [N1][N2]
    if ((st = row_dfts(fe, b0, b1, N2, N1, 0)) != FDAS_OK) return st;      // DFT over n2: Z[k1 + N1 k2]
    if ((st = transpose(fe, b1, b0, N1, N2)) != FDAS_OK) return st;       // natural order
    Z = b0;
  }
  const uint32_t grid = std::min<uint32_t>((N + 255) / 256, 4096);
  hipLaunchKernelGGL(real_spectrum_kernel, dim3(grid), dim3(256), 0, fe->stream, Z, X, N, log2n,
                     1.0f / float(2.0 * N));
  FEHIP(fe, hipGetLastError());
  if (!(flags & FDAS_FE_DEVICE_OUTPUT))
    FEHIP(fe, hipMemcpyAsync(spectrum_c64, X, sizeof(float2) * N, hipMemcpyDeviceToHost, fe->stream));
  FEHIP(fe, hipStreamSynchronize(fe->stream));
  return FDAS_OK;
}

}  // extern "C"
