/*
 * CPU restatement of the reference's FDAS path in C (test and baseline
 * infrastructure only; the GPU product never calls it).
 *
 *   fdas_cpu_fop   : FOP[t][f] = |sum_m h_t[m] X[f-m]|^2 (ft_conv.py:199-207), computed
 *                    by overlap-save with tile S and overlap L-1 (ft_conv.py:158-183),
 *                    own iterative radix-2 FFT, one template per OpenMP task.
 *   fdas_cpu_hsum  : HP_1 = FOP, HP_k = HP_{k-1} + FOP[trunc(t/k)][f/k] (hsum.py:91-120),
 *                    strict HP_k > thr_k detections (hsum.py:133-151), capped per plane
 *                    to the `cap` strongest (ties by (row, bin)) and sorted by (row, bin).
 *
 * REAL is float (the CPU baseline, same arithmetic type as the GPU) or double
 * (compiled as a second library for FP64 checks).  Pinned against the NumPy
 * oracle in tests/test_oracle_c.py.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef REAL
#define REAL float
#endif

typedef struct {
  REAL re, im;
} cpx;

static void fft_plan(int n, cpx* tw) {
  for (int k = 0; k < n / 2; ++k) {
    const double a = -2.0 * 3.14159265358979323846 * (double)k / (double)n;
    tw[k].re = (REAL)cos(a);
    tw[k].im = (REAL)sin(a);
  }
}

/* In-place iterative radix-2 DFT; dir = -1 forward, +1 inverse (unnormalised). */
static void fft_run(cpx* a, int n, const cpx* tw, int dir) {
  for (int i = 1, j = 0; i < n; ++i) {
    int bit = n >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j) {
      cpx t = a[i];
      a[i] = a[j];
      a[j] = t;
    }
  }
  for (int len = 2; len <= n; len <<= 1) {
    const int step = n / len;
    for (int i = 0; i < n; i += len) {
      for (int k = 0; k < len / 2; ++k) {
        cpx w = tw[k * step];
        if (dir > 0) w.im = -w.im;
        cpx u = a[i + k], v = a[i + k + len / 2];
        const REAL vr = v.re * w.re - v.im * w.im, vi = v.re * w.im + v.im * w.re;
        a[i + k].re = u.re + vr;
        a[i + k].im = u.im + vi;
        a[i + k + len / 2].re = u.re - vr;
        a[i + k + len / 2].im = u.im - vi;
      }
    }
  }
}

int fdas_cpu_fop(const float* x_c64, int n, const float* taps_c64, int t_count, int taps, int tile, float* fop,
                 int n_threads) {
  if (!x_c64 || !taps_c64 || !fop || n <= 0 || t_count <= 0 || taps <= 0 || tile < 2 * taps) return -1;
  const int o = taps - 1, p = tile - o;
  const int n_tiles = (n + p - 1) / p;
  cpx* tw = (cpx*)malloc(sizeof(cpx) * (size_t)tile / 2);
  cpx* xt = (cpx*)malloc(sizeof(cpx) * (size_t)tile * (size_t)n_tiles);
  if (!tw || !xt) {
    free(tw);
    free(xt);
    return -2;
  }
  fft_plan(tile, tw);
#ifdef _OPENMP
  if (n_threads > 0) omp_set_num_threads(n_threads);
#endif
  /* forward transform of every tile: tile j holds X[j*P - O + s] (ft_conv.py:164-180) */
#pragma omp parallel for schedule(static)
  for (int j = 0; j < n_tiles; ++j) {
    cpx* t = xt + (size_t)j * tile;
    for (int s = 0; s < tile; ++s) {
      const long idx = (long)j * p - o + s;
      if (idx >= 0 && idx < n) {
        t[s].re = (REAL)x_c64[2 * idx];
        t[s].im = (REAL)x_c64[2 * idx + 1];
      } else {
        t[s].re = t[s].im = 0;
      }
    }
    fft_run(t, tile, tw, -1);
  }
  int err = 0;
#pragma omp parallel
  {
    cpx* h = (cpx*)malloc(sizeof(cpx) * (size_t)tile);
    cpx* y = (cpx*)malloc(sizeof(cpx) * (size_t)tile);
    if (!h || !y) {
#pragma omp atomic write
      err = -2;
    }
#pragma omp for schedule(dynamic, 1)
    for (int r = 0; r < t_count; ++r) {
      if (!h || !y) continue;
      for (int s = 0; s < tile; ++s) {
        h[s].re = h[s].im = 0;
        if (s < taps) {
          h[s].re = (REAL)taps_c64[2 * ((size_t)r * taps + s)];
          h[s].im = (REAL)taps_c64[2 * ((size_t)r * taps + s) + 1];
        }
      }
      fft_run(h, tile, tw, -1);
      const REAL inv = (REAL)1 / (REAL)tile;
      for (int j = 0; j < n_tiles; ++j) {
        const cpx* t = xt + (size_t)j * tile;
        for (int s = 0; s < tile; ++s) {
          y[s].re = (t[s].re * h[s].re - t[s].im * h[s].im) * inv;
          y[s].im = (t[s].re * h[s].im + t[s].im * h[s].re) * inv;
        }
        fft_run(y, tile, tw, +1);
        for (int s = o; s < tile; ++s) {
          const long f = (long)j * p + s - o;
          if (f < n) fop[(size_t)r * n + f] = (float)(y[s].re * y[s].re + y[s].im * y[s].im);
        }
      }
    }
    free(h);
    free(y);
  }
  free(tw);
  free(xt);
  return err;
}

typedef struct {
  float pwr;
  uint32_t key;
} cand;

static int by_strength(const void* a, const void* b) {
  const cand* x = (const cand*)a;
  const cand* y = (const cand*)b;
  if (x->pwr != y->pwr) return x->pwr > y->pwr ? -1 : 1;
  return (x->key > y->key) - (x->key < y->key);
}

static int by_key(const void* a, const void* b) {
  const cand* x = (const cand*)a;
  const cand* y = (const cand*)b;
  return (x->key > y->key) - (x->key < y->key);
}

/* keys are row << 22 | bin, like the device lists */
int fdas_cpu_hsum(const float* fop, int t_count, int n, int n_harm, const float* thr, int cap, uint32_t* n_detected,
                  uint32_t* key_out, float* pwr_out, uint32_t* n_out, int n_threads) {
  if (!fop || !thr || n <= 0 || t_count <= 0 || n_harm <= 0 || n_harm > 16 || cap <= 0) return -1;
  const int c = t_count / 2;
#ifdef _OPENMP
  if (n_threads > 0) omp_set_num_threads(n_threads);
#endif
  int err = 0;
  for (int k = 0; k < n_harm; ++k) n_detected[k] = 0;
  /* per-row detection lists, merged per plane afterwards */
  cand** lists = (cand**)calloc((size_t)t_count * n_harm, sizeof(cand*));
  size_t* lens = (size_t*)calloc((size_t)t_count * n_harm, sizeof(size_t));
  if (!lists || !lens) {
    free(lists);
    free(lens);
    return -2;
  }
#pragma omp parallel
  {
    REAL* hp = (REAL*)malloc(sizeof(REAL) * (size_t)n);
#pragma omp for schedule(dynamic, 1)
    for (int r = 0; r < t_count; ++r) {
      if (!hp) {
#pragma omp atomic write
        err = -2;
        continue;
      }
      const int t = r - c;
      for (int k = 1; k <= n_harm; ++k) {
        const float* src = fop + (size_t)(c + t / k) * n; /* C division truncates toward zero */
        if (k == 1)
          for (int f = 0; f < n; ++f) hp[f] = (REAL)src[f];
        else
          for (int f = 0; f < n; ++f) hp[f] += (REAL)src[f / k];
        const REAL th = (REAL)thr[k - 1];
        size_t cnt = 0, capl = 0;
        cand* l = NULL;
        for (int f = 0; f < n; ++f) {
          if (hp[f] > th) {
            if (cnt == capl) {
              capl = capl ? 2 * capl : 64;
              cand* nl = (cand*)realloc(l, capl * sizeof(cand));
              if (!nl) {
#pragma omp atomic write
                err = -2;
                break;
              }
              l = nl;
            }
            l[cnt].pwr = (float)hp[f];
            l[cnt].key = ((uint32_t)r << 22) | (uint32_t)f;
            ++cnt;
          }
        }
        lists[(size_t)r * n_harm + (k - 1)] = l;
        lens[(size_t)r * n_harm + (k - 1)] = cnt;
      }
    }
    free(hp);
  }
  for (int k = 0; k < n_harm && !err; ++k) {
    size_t total = 0;
    for (int r = 0; r < t_count; ++r) total += lens[(size_t)r * n_harm + k];
    cand* all = (cand*)malloc(sizeof(cand) * (total ? total : 1));
    if (!all) {
      err = -2;
      break;
    }
    size_t w = 0;
    for (int r = 0; r < t_count; ++r) {
      const size_t m = lens[(size_t)r * n_harm + k];
      if (m) memcpy(all + w, lists[(size_t)r * n_harm + k], m * sizeof(cand));
      w += m;
    }
    n_detected[k] = (uint32_t)total;
    size_t keep = total;
    if (total > (size_t)cap) {
      qsort(all, total, sizeof(cand), by_strength);
      keep = (size_t)cap;
    }
    qsort(all, keep, sizeof(cand), by_key);
    for (size_t i = 0; i < keep; ++i) {
      key_out[(size_t)k * cap + i] = all[i].key;
      pwr_out[(size_t)k * cap + i] = all[i].pwr;
    }
    n_out[k] = (uint32_t)keep;
    free(all);
  }
  for (size_t i = 0; i < (size_t)t_count * n_harm; ++i) free(lists[i]);
  free(lists);
  free(lens);
  return err;
}

int fdas_cpu_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
