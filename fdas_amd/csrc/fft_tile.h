// Register/LDS FFT building blocks for the fused FT-convolution kernel (gfx950).
//
// A tile of S = 16 * NT complex points is transformed by NT = 16 * D threads
// (D = 4, 8, 16 for S = 1024, 2048, 4096), 16 points per thread, in three
// register passes separated by two LDS exchanges:
//
//   F1  thread b        : DFT-16 over a of x[NT*a + b], twiddle W_S^(b*ka)
//   F2  thread (d, ka)  : DFT-16 over c of Y[D*c + d][ka], twiddle W_NT^(d*kc)
//   F3  pairs q = 16*kc + ka : DFT-D over d
//
// giving X[ka + 16*kc + 256*kd] = X[q + 256*kd].  The inverse runs the same
// passes in reverse order with conjugate twiddles, so the spectrum never needs
// a digit-reversal permutation: the template spectra are stored in the order
// the F3 pass leaves the tile in (see ftconv_slot_bin).
//
// The unnormalised DFT matches the reference's FFT engine (fft_4p.cl:228-319),
// whose inverse is swap(re,im) o DFT o swap(re,im) (fft_4p.cl:77-88).
#pragma once

#include <hip/hip_runtime.h>

namespace fdas {

// Global-memory access helpers.  Pointers read from a job descriptor are
// generic to the compiler, which then emits FLAT instructions; FLAT memory ops
// also count against the LDS wait counter, so every LDS wait would drain the
// outstanding HBM traffic too.  These helpers address global memory (address
// space 1) explicitly, so the loads and stores stay global_load / global_store.
#define FDAS_GLOBAL __attribute__((address_space(1)))
typedef float fdas_vf2 __attribute__((ext_vector_type(2)));
typedef float fdas_vf4 __attribute__((ext_vector_type(4)));

template <typename T>
__device__ __forceinline__ T gld(const T* p, size_t i) {
  return ((const FDAS_GLOBAL T*)p)[i];
}
template <typename T>
__device__ __forceinline__ void gst(T* p, size_t i, T v) {
  ((FDAS_GLOBAL T*)p)[i] = v;
}
__device__ __forceinline__ float2 gld2(const float2* p, size_t i) {
  const fdas_vf2 v = ((const FDAS_GLOBAL fdas_vf2*)p)[i];
  return make_float2(v.x, v.y);
}
// Four consecutive floats p[i .. i+3]; p + i must be 16-byte aligned.
__device__ __forceinline__ fdas_vf4 gld4(const float* p, size_t i) {
  return *(const FDAS_GLOBAL fdas_vf4*)(p + i);
}
// Streaming store: the data is read back only by a later kernel, so it is
// written with the non-temporal hint and does not displace the cached
// template spectra and twiddles.
__device__ __forceinline__ void gst_stream(float* p, size_t i, float v) {
  __builtin_nontemporal_store(v, (FDAS_GLOBAL float*)p + i);
}
__device__ __forceinline__ uint32_t gatomic_add(uint32_t* p, uint32_t v) {
  return __hip_atomic_fetch_add((FDAS_GLOBAL uint32_t*)p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * conj(b)
__device__ __forceinline__ float2 cmulc(float2 a, float2 b) {
  return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
// a * W with W = conj(w) for the inverse direction (DIR = +1), w for DIR = -1.
template <int DIR>
__device__ __forceinline__ float2 ctw(float2 a, float2 w) {
  if constexpr (DIR < 0) return cmul(a, w);
  else return cmulc(a, w);
}

// cos / sin of e*pi/8, e = 0..15.
struct Oct16 {
  static constexpr double c[16] = {1.0, 0.92387953251128674, 0.70710678118654752, 0.38268343236508978,
                                   0.0, -0.38268343236508978, -0.70710678118654752, -0.92387953251128674,
                                   -1.0, -0.92387953251128674, -0.70710678118654752, -0.38268343236508978,
                                   0.0, 0.38268343236508978, 0.70710678118654752, 0.92387953251128674};
  static constexpr double s[16] = {0.0, 0.38268343236508978, 0.70710678118654752, 0.92387953251128674,
                                   1.0, 0.92387953251128674, 0.70710678118654752, 0.38268343236508978,
                                   0.0, -0.38268343236508978, -0.70710678118654752, -0.92387953251128674,
                                   -1.0, -0.92387953251128674, -0.70710678118654752, -0.38268343236508978};
};

// a * exp(DIR * 2*pi*i * E / M) for M in {4, 8, 16}; special angles cost no multiply.
template <int DIR, int M, int E>
__device__ __forceinline__ float2 rot(float2 a) {
  static_assert(M == 4 || M == 8 || M == 16, "rot: M must divide 16");
  constexpr int e16 = (((E % M) + M) % M) * (16 / M);  // angle in units of pi/8
  constexpr float r = 0.70710678118654752440f;
  if constexpr (e16 == 0) {
    return a;
  } else if constexpr (e16 == 8) {
    return make_float2(-a.x, -a.y);
  } else if constexpr (e16 == 4) {  // * DIR*i
    return DIR < 0 ? make_float2(a.y, -a.x) : make_float2(-a.y, a.x);
  } else if constexpr (e16 == 12) {  // * -DIR*i
    return DIR < 0 ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
  } else if constexpr (e16 == 2) {  // (1 + DIR i)/sqrt2
    return DIR < 0 ? make_float2((a.x + a.y) * r, (a.y - a.x) * r) : make_float2((a.x - a.y) * r, (a.y + a.x) * r);
  } else if constexpr (e16 == 6) {  // (-1 + DIR i)/sqrt2
    return DIR < 0 ? make_float2((a.y - a.x) * r, (-a.y - a.x) * r) : make_float2((-a.x - a.y) * r, (a.x - a.y) * r);
  } else if constexpr (e16 == 10) {  // (-1 - DIR i)/sqrt2
    return DIR < 0 ? make_float2((-a.x - a.y) * r, (a.x - a.y) * r) : make_float2((a.y - a.x) * r, (-a.y - a.x) * r);
  } else if constexpr (e16 == 14) {  // (1 - DIR i)/sqrt2
    return DIR < 0 ? make_float2((a.x - a.y) * r, (a.y + a.x) * r) : make_float2((a.x + a.y) * r, (a.y - a.x) * r);
  } else {
    constexpr float cc = static_cast<float>(Oct16::c[e16]);
    constexpr float ss = static_cast<float>(DIR * Oct16::s[e16]);
    return make_float2(a.x * cc - a.y * ss, a.x * ss + a.y * cc);
  }
}

// In-place DFT-4 of v[o], v[o+st], v[o+2st], v[o+3st]  (outputs in natural order).
template <int DIR, int O, int ST, int N>
__device__ __forceinline__ void dft4(float2 (&v)[N]) {
  const float2 a0 = v[O], a1 = v[O + ST], a2 = v[O + 2 * ST], a3 = v[O + 3 * ST];
  const float2 t0 = cadd(a0, a2), t1 = csub(a0, a2);
  const float2 t2 = cadd(a1, a3), t3 = rot<DIR, 4, 1>(csub(a1, a3));
  v[O] = cadd(t0, t2);
  v[O + 2 * ST] = csub(t0, t2);
  v[O + ST] = cadd(t1, t3);
  v[O + 3 * ST] = csub(t1, t3);
}

// DFT-8 of v[o + st*n], n = 0..7, natural-order output.
template <int DIR, int O, int ST, int N>
__device__ __forceinline__ void dft8(float2 (&v)[N]) {
  float2 e[4] = {v[O], v[O + 2 * ST], v[O + 4 * ST], v[O + 6 * ST]};
  float2 d[4] = {v[O + ST], v[O + 3 * ST], v[O + 5 * ST], v[O + 7 * ST]};
  dft4<DIR, 0, 1>(e);
  dft4<DIR, 0, 1>(d);
  d[1] = rot<DIR, 8, 1>(d[1]);
  d[2] = rot<DIR, 8, 2>(d[2]);
  d[3] = rot<DIR, 8, 3>(d[3]);
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    v[O + k * ST] = cadd(e[k], d[k]);
    v[O + (k + 4) * ST] = csub(e[k], d[k]);
  }
}

template <int DIR, int K1, int N2>
__device__ __forceinline__ float2 tw16(float2 a) { return rot<DIR, 16, K1 * N2>(a); }

// DFT-4 over a[n2] = v[4 K1 + n2] twiddled by exp(DIR 2 pi i n2 K1 / 16), K1 = 1..3, in place
// (natural order).  Each twiddle cos + i DIR sin is applied as cos * (1 + i DIR tan): the odd
// inputs' cosines share one factor g (cos 3pi/8 = sin pi/8), which, like the 1/sqrt2 of the even
// input, rides on the butterfly's closing FMAs.  K1 = 1 and 3 take 22 VALU, K1 = 2 takes 20,
// against 28 and 24 for the twiddle multiplies followed by a DFT-4.
template <int DIR, int K1>
__device__ __forceinline__ void dft4_tw16(float2 (&v)[16]) {
  static_assert(K1 >= 1 && K1 <= 3, "row 0 has no twiddles");
  constexpr float D = float(DIR);
  constexpr float r = 0.70710678118654752440f;   // cos pi/4
  constexpr float t1 = 0.41421356237309504880f;  // tan pi/8
  constexpr float t3 = 2.41421356237309504880f;  // tan 3pi/8
  const float2 a0 = v[4 * K1], a1 = v[4 * K1 + 1], a2 = v[4 * K1 + 2], a3 = v[4 * K1 + 3];
  float2 t0, tl, u, w;  // a0 +- a2 w2 as t0 / tl; (a1 w1 +- a3 w3) / g as u / w
  float g;
  if constexpr (K1 == 2) {
    // w1 = r (1 + D i), w2 = D i, w3 = r (-1 + D i)
    t0 = make_float2(a0.x - D * a2.y, a0.y + D * a2.x);
    tl = make_float2(a0.x + D * a2.y, a0.y - D * a2.x);
    const float2 d = csub(a1, a3), s = cadd(a1, a3);
    u = make_float2(d.x - D * s.y, d.y + D * s.x);
    w = make_float2(s.x - D * d.y, s.y + D * d.x);
    g = r;
  } else {
    // K1 = 1: w1 = c1 (1 + D i t1), w2 = r (1 + D i), w3 = s1 (1 + D i t3), g = c1, w3 / g = t1 (1 + D i t3)
    // K1 = 3: w1 = s1 (1 + D i t3), w2 = r (-1 + D i), w3 = -c1 (1 + D i t1), g = s1, w3 / g = -t3 (1 + D i t1)
    constexpr float ta = K1 == 1 ? t1 : t3, tb = K1 == 1 ? t3 : t1, h = K1 == 1 ? t1 : -t3;
    g = K1 == 1 ? 0.92387953251128674f : 0.38268343236508978f;  // cos pi/8 | sin pi/8
    const float2 q1 = make_float2(fmaf(-D * ta, a1.y, a1.x), fmaf(D * ta, a1.x, a1.y));
    const float2 q3 = make_float2(fmaf(-D * tb, a3.y, a3.x), fmaf(D * tb, a3.x, a3.y));
    u = make_float2(fmaf(h, q3.x, q1.x), fmaf(h, q3.y, q1.y));
    w = make_float2(fmaf(-h, q3.x, q1.x), fmaf(-h, q3.y, q1.y));
    const float2 p = K1 == 1 ? make_float2(a2.x - D * a2.y, a2.y + D * a2.x)
                             : make_float2(-a2.x - D * a2.y, -a2.y + D * a2.x);  // a2 w2 / r
    t0 = make_float2(fmaf(r, p.x, a0.x), fmaf(r, p.y, a0.y));
    tl = make_float2(fmaf(-r, p.x, a0.x), fmaf(-r, p.y, a0.y));
  }
  // X0, X2 = t0 +- g u; X1, X3 = tl +- g (D i) w
  v[4 * K1] = make_float2(fmaf(g, u.x, t0.x), fmaf(g, u.y, t0.y));
  v[4 * K1 + 2] = make_float2(fmaf(-g, u.x, t0.x), fmaf(-g, u.y, t0.y));
  v[4 * K1 + 1] = make_float2(fmaf(-D * g, w.y, tl.x), fmaf(D * g, w.x, tl.y));
  v[4 * K1 + 3] = make_float2(fmaf(D * g, w.y, tl.x), fmaf(-D * g, w.x, tl.y));
}

// DFT-16 of v[0..15] (contiguous), natural-order output, as 4 x 4.
template <int DIR>
__device__ __forceinline__ void dft16(float2 (&v)[16]) {
  // Step 1: DFT-4 over n1 for each n2 (elements n2, 4+n2, 8+n2, 12+n2).
  dft4<DIR, 0, 4>(v);
  dft4<DIR, 1, 4>(v);
  dft4<DIR, 2, 4>(v);
  dft4<DIR, 3, 4>(v);
  // Now v[4*k1 + n2] holds B[n2][k1].  Step 2: twiddle by W16^(n2*k1), then DFT-4 over n2
  // for each k1: elements 4*k1 .. 4*k1+3 -> X[k1 + 4*k2].
  dft4<DIR, 0, 1>(v);
  dft4_tw16<DIR, 1>(v);
  dft4_tw16<DIR, 2>(v);
  dft4_tw16<DIR, 3>(v);
  // v[4*k1 + k2] = X[k1 + 4*k2]: transpose 4x4 in registers (renaming only).
  float2 t[16];
#pragma unroll
  for (int k1 = 0; k1 < 4; ++k1)
#pragma unroll
    for (int k2 = 0; k2 < 4; ++k2) t[k1 + 4 * k2] = v[4 * k1 + k2];
#pragma unroll
  for (int i = 0; i < 16; ++i) v[i] = t[i];
}

// DFT-D over v[O + i], i = 0..D-1 (contiguous group), natural-order output.
template <int DIR, int D, int O, int N>
__device__ __forceinline__ void dftD(float2 (&v)[N]) {
  if constexpr (D == 4) {
    dft4<DIR, O, 1>(v);
  } else if constexpr (D == 8) {
    dft8<DIR, O, 1>(v);
  } else {
    static_assert(D == 16, "D must be 4, 8 or 16");
    float2 t[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) t[i] = v[O + i];
    dft16<DIR>(t);
#pragma unroll
    for (int i = 0; i < 16; ++i) v[O + i] = t[i];
  }
}

// LDS row padding of the two exchanges (float2 units), chosen by measurement
// of LDS bank conflicts on gfx950.
#ifndef FDAS_RS1_PAD
#define FDAS_RS1_PAD 1
#endif
#ifndef FDAS_RS2_PAD
#define FDAS_RS2_PAD 16
#endif

// Compile-time layout of one S-point tile transform.
template <int S>
struct TileFFT {
  static_assert(S == 1024 || S == 2048 || S == 4096, "tile must be 1024, 2048 or 4096");
  static constexpr int NT = S / 16;     // threads per transform
  static constexpr int D = NT / 16;     // radix of the last pass
  static constexpr int PAIRS = 16 / D;  // (ka, kc) pairs per thread in pass 3
  static constexpr int RS1 = NT + FDAS_RS1_PAD;    // row stride of the [ka][b] exchange
  static constexpr int RS2 = 256 + FDAS_RS2_PAD;   // row stride of the [d][q] exchange
  static constexpr int L1_ELEMS = 16 * RS1;
  static constexpr int L2_ELEMS = D * RS2;
  static constexpr int LDS_ELEMS = L1_ELEMS > L2_ELEMS ? L1_ELEMS : L2_ELEMS;

  // Frequency bin held by thread `tid` in register slot `slot` after pass 3
  // (slot = i*D + kd for pair q = tid + NT*i).
  __host__ __device__ static constexpr int slot_bin(int tid, int slot) {
    return tid + NT * (slot / D) + 256 * (slot % D);
  }

  // Per-thread twiddles: w1[ka] = W_S^(b*ka) (b = tid), w2[kc] = W_NT^(d*kc) (d = tid/16).
  float2 w1[16];
  float2 w2[16];

  __device__ __forceinline__ void load_twiddles(const float2* __restrict__ tw1, const float2* __restrict__ tw2,
                                                int tid) {
    w1[0] = make_float2(1.f, 0.f);
    w2[0] = make_float2(1.f, 0.f);
#pragma unroll
    for (int k = 1; k < 16; ++k) {
      w1[k] = tw1[k * NT + tid];
      w2[k] = tw2[k * D + (tid >> 4)];
    }
  }

  // Forward transform: x (16 points, x[a] = tile[NT*a + tid]) -> v (16 spectrum
  // slots, slot_bin order).  Ends with a barrier after the last LDS read.
  __device__ __forceinline__ void forward(float2 (&v)[16], float2* __restrict__ lds, int tid) const {
    // F1
    dft16<-1>(v);
#pragma unroll
    for (int k = 1; k < 16; ++k) v[k] = cmul(v[k], w1[k]);
#pragma unroll
    for (int k = 0; k < 16; ++k) lds[k * RS1 + tid] = v[k];
    __syncthreads();
    // F2: thread (d, ka) reads Y[D*c + d][ka] for c = 0..15
    const int ka = tid & 15, d = tid >> 4;
#pragma unroll
    for (int c = 0; c < 16; ++c) v[c] = lds[ka * RS1 + D * c + d];
    __syncthreads();
    dft16<-1>(v);
#pragma unroll
    for (int k = 1; k < 16; ++k) v[k] = cmul(v[k], w2[k]);
#pragma unroll
    for (int kc = 0; kc < 16; ++kc) lds[d * RS2 + 16 * kc + ka] = v[kc];
    __syncthreads();
    // F3: pairs q = tid + NT*i, DFT-D over d
#pragma unroll
    for (int i = 0; i < PAIRS; ++i)
#pragma unroll
      for (int dd = 0; dd < D; ++dd) v[i * D + dd] = lds[dd * RS2 + tid + NT * i];
    __syncthreads();
    apply_dftD<-1>(v);
  }

  // Inverse transform of spectrum slots v -> v[a] = out[NT*a + tid] (unnormalised).
  // Uses two LDS buffers: a (pair layout) and b (row layout).  The caller must
  // ensure no thread still reads `a` from a previous forward pass.
  __device__ __forceinline__ void inverse(float2 (&v)[16], float2* __restrict__ la, float2* __restrict__ lb,
                                          int tid) const {
    apply_dftD<+1>(v);
#pragma unroll
    for (int i = 0; i < PAIRS; ++i)
#pragma unroll
      for (int dd = 0; dd < D; ++dd) la[dd * RS2 + tid + NT * i] = v[i * D + dd];
    __syncthreads();
    const int ka = tid & 15, d = tid >> 4;
#pragma unroll
    for (int kc = 0; kc < 16; ++kc) v[kc] = la[d * RS2 + 16 * kc + ka];
#pragma unroll
    for (int k = 1; k < 16; ++k) v[k] = cmulc(v[k], w2[k]);
    dft16<+1>(v);
#pragma unroll
    for (int c = 0; c < 16; ++c) lb[ka * RS1 + D * c + d] = v[c];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 16; ++k) v[k] = lb[k * RS1 + tid];
#pragma unroll
    for (int k = 1; k < 16; ++k) v[k] = cmulc(v[k], w1[k]);
    dft16<+1>(v);
  }

  // As inverse(), with one LDS buffer for both exchanges (half the LDS, two more
  // barriers).  Starts with a barrier, so the caller's previous readers of `lds`
  // are done before it is rewritten.
  __device__ __forceinline__ void inverse_single(float2 (&v)[16], float2* __restrict__ lds, int tid) const {
    apply_dftD<+1>(v);
    __syncthreads();
#pragma unroll
    for (int i = 0; i < PAIRS; ++i)
#pragma unroll
      for (int dd = 0; dd < D; ++dd) lds[dd * RS2 + tid + NT * i] = v[i * D + dd];
    __syncthreads();
    const int ka = tid & 15, d = tid >> 4;
#pragma unroll
    for (int kc = 0; kc < 16; ++kc) v[kc] = lds[d * RS2 + 16 * kc + ka];
#pragma unroll
    for (int k = 1; k < 16; ++k) v[k] = cmulc(v[k], w2[k]);
    dft16<+1>(v);
    __syncthreads();
#pragma unroll
    for (int c = 0; c < 16; ++c) lds[ka * RS1 + D * c + d] = v[c];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 16; ++k) v[k] = lds[k * RS1 + tid];
#pragma unroll
    for (int k = 1; k < 16; ++k) v[k] = cmulc(v[k], w1[k]);
    dft16<+1>(v);
  }

 private:
  template <int DIR>
  __device__ __forceinline__ static void apply_dftD(float2 (&v)[16]) {
    if constexpr (PAIRS == 1) {
      dftD<DIR, D, 0>(v);
    } else if constexpr (PAIRS == 2) {
      dftD<DIR, D, 0>(v);
      dftD<DIR, D, D>(v);
    } else {
      dftD<DIR, D, 0>(v);
      dftD<DIR, D, D>(v);
      dftD<DIR, D, 2 * D>(v);
      dftD<DIR, D, 3 * D>(v);
    }
  }
};


// ---------------------------------------------------------------------------
// One-wave 1024-point transform whose first exchange stays in registers.
//
// The same three passes as TileFFT<1024> (F1 DFT-16, F2 DFT-16, F3 DFT-4 and
// the mirror-image inverse), with the same spectrum slot order, so template
// spectra laid out for TileFFT<1024> serve both.  The exchange between F2 and
// F3 only swaps the two high lane bits (lane / 16) with the two low register
// bits (slot mod 4): thread (ka, m) slot 4i + dd <-> thread (ka, dd) slot
// 4i + m, ka = lane mod 16.  gfx950's v_permlane16_swap (lane bit 4) and
// v_permlane32_swap (lane bit 5) do each 2x2 block transpose in one
// instruction per dword, without LDS.  The other exchange (F1 <-> F2) moves the
// low lane bits and goes through the wave's own LDS buffer; it needs no
// workgroup barrier, since only this wave reads what it wrote.
// ---------------------------------------------------------------------------

// 2x2 transpose of (lane bit 4, register pair): x holds the bit-0 register,
// y the bit-1 register.  Odd 16-lane rows of x trade places with even rows of y.
__device__ __forceinline__ void swap_lane_bit4(float2& x, float2& y) {
  const auto re = __builtin_amdgcn_permlane16_swap(__float_as_uint(x.x), __float_as_uint(y.x), false, false);
  const auto im = __builtin_amdgcn_permlane16_swap(__float_as_uint(x.y), __float_as_uint(y.y), false, false);
  x = make_float2(__uint_as_float(re[0]), __uint_as_float(im[0]));
  y = make_float2(__uint_as_float(re[1]), __uint_as_float(im[1]));
}
// The same for lane bit 5: the upper 32 lanes of x trade places with the lower 32 of y.
__device__ __forceinline__ void swap_lane_bit5(float2& x, float2& y) {
  const auto re = __builtin_amdgcn_permlane32_swap(__float_as_uint(x.x), __float_as_uint(y.x), false, false);
  const auto im = __builtin_amdgcn_permlane32_swap(__float_as_uint(x.y), __float_as_uint(y.y), false, false);
  x = make_float2(__uint_as_float(re[0]), __uint_as_float(im[0]));
  y = make_float2(__uint_as_float(re[1]), __uint_as_float(im[1]));
}

// Thread (ka, m) slot 4i + dd <-> thread (ka, dd) slot 4i + m (an involution).
__device__ __forceinline__ void swap_lane_hi_slot_lo(float2 (&v)[16]) {
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    swap_lane_bit4(v[4 * i + 0], v[4 * i + 1]);
    swap_lane_bit4(v[4 * i + 2], v[4 * i + 3]);
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    swap_lane_bit5(v[4 * i + 0], v[4 * i + 2]);
    swap_lane_bit5(v[4 * i + 1], v[4 * i + 3]);
  }
}

// Orders this wave's LDS accesses: the compiler may not move loads or stores
// across it, and one wave's LDS instructions complete in issue order.
__device__ __forceinline__ void wave_lds_order() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// The F2 twiddles W_64^(d * kc) (d = lane / 16) are read from a 64-entry LDS
// table shared by the workgroup (15 fewer complex registers per thread).
template <bool W2_LDS = true>
struct WaveFFT1k {
  static_assert(W2_LDS, "the F2 twiddles of this form live in LDS");
  static constexpr int S = 1024;
  static constexpr int NT = 64;
  static constexpr int RS1 = NT + FDAS_RS1_PAD;  // row stride of the [ka][b] exchange
  static constexpr int LDS_ELEMS = 16 * RS1;     // one wave's exchange buffer (float2)

  float2 w1[16];                     // W_1024^(lane * ka)
  float2 w2r[W2_LDS ? 1 : 16];       // W_64^(d * kc), registers
  const float2* __restrict__ w2;     // or its LDS copy at [kc * 4 + d]

  // w2_lds: 64 float2 of LDS the calling wave fills (lanes 0..63 copy one entry
  // each) when W2_LDS; it is read after the caller's next barrier.
  __device__ __forceinline__ void load_twiddles(const float2* __restrict__ tw1, const float2* __restrict__ tw2,
                                                float2* w2_lds, int lane, bool fill) {
    w1[0] = make_float2(1.f, 0.f);
#pragma unroll
    for (int k = 1; k < 16; ++k) w1[k] = tw1[k * NT + lane];
    if constexpr (W2_LDS) {
      if (fill) w2_lds[lane] = tw2[lane];
      w2 = w2_lds;
    } else {
      w2r[0] = make_float2(1.f, 0.f);
#pragma unroll
      for (int k = 1; k < 16; ++k) w2r[k] = tw2[k * 4 + (lane >> 4)];
      w2 = nullptr;
    }
  }

  __device__ __forceinline__ float2 tw2(int k, int d) const {
    if constexpr (W2_LDS) return w2[k * 4 + d];
    else return w2r[k];
  }

  // x[a] = tile[64 a + lane] -> spectrum slots (TileFFT<1024>::slot_bin order).
  __device__ __forceinline__ void forward(float2 (&v)[16], float2* __restrict__ lb, int lane) const {
    dft16<-1>(v);
#pragma unroll
    for (int k = 1; k < 16; ++k) v[k] = cmul(v[k], w1[k]);
    wave_lds_order();
#pragma unroll
    for (int k = 0; k < 16; ++k) lb[k * RS1 + lane] = v[k];
    wave_lds_order();
    const int ka = lane & 15, d = lane >> 4;
#pragma unroll
    for (int c = 0; c < 16; ++c) v[c] = lb[ka * RS1 + 4 * c + d];
    dft16<-1>(v);
#pragma unroll
    for (int k = 1; k < 16; ++k) v[k] = cmul(v[k], tw2(k, d));
    swap_lane_hi_slot_lo(v);
    dft4<-1, 0, 1>(v);
    dft4<-1, 4, 1>(v);
    dft4<-1, 8, 1>(v);
    dft4<-1, 12, 1>(v);
  }

  // Spectrum slots -> v[a] = out[64 a + lane] (unnormalised inverse).
  __device__ __forceinline__ void inverse(float2 (&v)[16], float2* __restrict__ lb, int lane) const {
    dft4<+1, 0, 1>(v);
    dft4<+1, 4, 1>(v);
    dft4<+1, 8, 1>(v);
    dft4<+1, 12, 1>(v);
    swap_lane_hi_slot_lo(v);
    const int ka = lane & 15, d = lane >> 4;
#pragma unroll
    for (int k = 1; k < 16; ++k) v[k] = cmulc(v[k], tw2(k, d));
    dft16<+1>(v);
    wave_lds_order();
#pragma unroll
    for (int c = 0; c < 16; ++c) lb[ka * RS1 + 4 * c + d] = v[c];
    wave_lds_order();
#pragma unroll
    for (int k = 0; k < 16; ++k) v[k] = lb[k * RS1 + lane];
#pragma unroll
    for (int k = 1; k < 16; ++k) v[k] = cmulc(v[k], w1[k]);
    dft16<+1>(v);
  }
};

// The one-wave transform with fewer live registers: the F2 twiddles are
// applied on the other side of the lane exchange, where register 4i + dd of
// lane (ka, m) needs W_64^(dd (4i + m)) = W_16^(dd i) W_64^(dd m): a
// compile-time rotation times one of three per-lane values u_dd = W_64^(dd m),
// m = lane / 16 (6 registers instead of 30; registers with dd = 0 need none).
struct WaveFFT1kLean {
  static constexpr int S = 1024;
  static constexpr int NT = 64;
  static constexpr int RS1 = NT + FDAS_RS1_PAD;
  static constexpr int LDS_ELEMS = 16 * RS1;

  float2 w1[16];  // W_1024^(lane * ka)
  float2 u[4];    // u[dd] = W_64^(dd * (lane / 16)), dd = 1..3

  __device__ __forceinline__ void load_twiddles(const float2* __restrict__ tw1, const float2* __restrict__ tw2,
                                                int lane) {
    w1[0] = make_float2(1.f, 0.f);
#pragma unroll
    for (int k = 1; k < 16; ++k) w1[k] = tw1[k * NT + lane];
    u[0] = make_float2(1.f, 0.f);
#pragma unroll
    for (int dd = 1; dd < 4; ++dd) u[dd] = tw2[dd * 4 + (lane >> 4)];  // tw2[k * 4 + d] = W_64^(d k)
  }

  // v[4i + dd] *= (W_16^(dd i) u_dd) for DIR = -1, its conjugate for DIR = +1.
  template <int DIR, int I, int DD>
  __device__ __forceinline__ void tw_one(float2 (&v)[16]) const {
    const float2 t = DIR < 0 ? cmul(v[4 * I + DD], u[DD]) : cmulc(v[4 * I + DD], u[DD]);
    v[4 * I + DD] = rot<DIR, 16, DD * I>(t);
  }
  template <int DIR, int I>
  __device__ __forceinline__ void tw_row(float2 (&v)[16]) const {
    tw_one<DIR, I, 1>(v);
    tw_one<DIR, I, 2>(v);
    tw_one<DIR, I, 3>(v);
  }
  template <int DIR>
  __device__ __forceinline__ void twiddle_f2(float2 (&v)[16]) const {
    tw_row<DIR, 0>(v);
    tw_row<DIR, 1>(v);
    tw_row<DIR, 2>(v);
    tw_row<DIR, 3>(v);
  }

  __device__ __forceinline__ void forward(float2 (&v)[16], float2* __restrict__ lb, int lane) const {
    dft16<-1>(v);
#pragma unroll
    for (int k = 1; k < 16; ++k) v[k] = cmul(v[k], w1[k]);
    wave_lds_order();
#pragma unroll
    for (int k = 0; k < 16; ++k) lb[k * RS1 + lane] = v[k];
    wave_lds_order();
    const int ka = lane & 15, d = lane >> 4;
#pragma unroll
    for (int c = 0; c < 16; ++c) v[c] = lb[ka * RS1 + 4 * c + d];
    dft16<-1>(v);
    swap_lane_hi_slot_lo(v);
    twiddle_f2<-1>(v);
    dft4<-1, 0, 1>(v);
    dft4<-1, 4, 1>(v);
    dft4<-1, 8, 1>(v);
    dft4<-1, 12, 1>(v);
  }

  __device__ __forceinline__ void inverse(float2 (&v)[16], float2* __restrict__ lb, int lane) const {
    dft4<+1, 0, 1>(v);
    dft4<+1, 4, 1>(v);
    dft4<+1, 8, 1>(v);
    dft4<+1, 12, 1>(v);
    twiddle_f2<+1>(v);
    swap_lane_hi_slot_lo(v);
    dft16<+1>(v);
    const int ka = lane & 15, d = lane >> 4;
    wave_lds_order();
#pragma unroll
    for (int c = 0; c < 16; ++c) lb[ka * RS1 + 4 * c + d] = v[c];
    wave_lds_order();
#pragma unroll
    for (int k = 0; k < 16; ++k) v[k] = lb[k * RS1 + lane];
#pragma unroll
    for (int k = 1; k < 16; ++k) v[k] = cmulc(v[k], w1[k]);
    dft16<+1>(v);
  }
};

}  // namespace fdas
