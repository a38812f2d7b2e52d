/*
 * CPU restatement of the reference's FDAS path in C (test and baseline
 * infrastructure only; the GPU product never calls it).
 *
 *   fdas_cpu_fop   : FOP[t][f] = |sum_m h_t[m] X[f-m]|^2 (ft_conv.py:199-207), computed
 *                    by overlap-save with tile S and overlap L-1 (ft_conv.py:158-183),
 *                    own self-sorting radix-4 FFT on split re/im arrays (vectorised
 *                    inner loops), one template per OpenMP task.
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

/* Self-sorting (Stockham) radix-4 FFT on split re/im arrays, decimation in
 * frequency, with a final radix-2 stage when log2(n) is odd.  Stage t has
 * stride s = 4^t and length n_t = n / 4^t; its inner loop runs over s
 * consecutive elements, which the compiler vectorises once s >= 8.  Twiddles
 * W_n_t^p, ^2p, ^3p of every stage are tabulated in FP64 and rounded once. */
typedef struct {
  int n, n_stages;  /* radix-4 stages; a radix-2 stage follows when n_stages' lengths end at 2 */
  REAL* tw;         /* per stage, m = n_t / 4 entries of (w1, w2, w3) as 6 REALs */
} fft_plan;

static int fft_plan_make(fft_plan* pl, int n) {
  pl->n = n;
  pl->n_stages = 0;
  size_t total = 0;
  for (int nt = n; nt >= 4; nt /= 4) {
    total += (size_t)(nt / 4) * 6;
    pl->n_stages++;
  }
  pl->tw = (REAL*)malloc(sizeof(REAL) * (total ? total : 1));
  if (!pl->tw) return -2;
  REAL* w = pl->tw;
  for (int nt = n; nt >= 4; nt /= 4) {
    const int m = nt / 4;
    for (int p = 0; p < m; ++p) {
      for (int e = 1; e <= 3; ++e) {
        const double a = -2.0 * 3.14159265358979323846 * (double)(((long)e * p) % nt) / (double)nt;
        w[6 * p + 2 * (e - 1)] = (REAL)cos(a);
        w[6 * p + 2 * (e - 1) + 1] = (REAL)sin(a);
      }
    }
    w += (size_t)m * 6;
  }
  return 0;
}

static void fft_plan_free(fft_plan* pl) { free(pl->tw); }

/* Forward DFT (e^{-2 pi i f n / N}, unnormalised) of (xr, xi) in place; (yr, yi) is work space. */
static void fft_forward(const fft_plan* pl, REAL* xr, REAL* xi, REAL* yr, REAL* yi) {
  const int n = pl->n;
  REAL *ar = xr, *ai = xi, *br = yr, *bi = yi;
  const REAL* w = pl->tw;
  int s = 1, nt = n;
  for (; nt >= 4; nt /= 4, s *= 4) {
    const int m = nt / 4;
    for (int p = 0; p < m; ++p) {
      const REAL w1r = w[6 * p], w1i = w[6 * p + 1], w2r = w[6 * p + 2], w2i = w[6 * p + 3];
      const REAL w3r = w[6 * p + 4], w3i = w[6 * p + 5];
      const REAL *a0r = ar + (size_t)s * p, *a0i = ai + (size_t)s * p;
      const REAL *a1r = a0r + (size_t)s * m, *a1i = a0i + (size_t)s * m;
      const REAL *a2r = a1r + (size_t)s * m, *a2i = a1i + (size_t)s * m;
      const REAL *a3r = a2r + (size_t)s * m, *a3i = a2i + (size_t)s * m;
      REAL *y0r = br + (size_t)s * 4 * p, *y0i = bi + (size_t)s * 4 * p;
      REAL *y1r = y0r + s, *y1i = y0i + s, *y2r = y1r + s, *y2i = y1i + s, *y3r = y2r + s, *y3i = y2i + s;
      for (int q = 0; q < s; ++q) {
        const REAL apcr = a0r[q] + a2r[q], apci = a0i[q] + a2i[q];
        const REAL amcr = a0r[q] - a2r[q], amci = a0i[q] - a2i[q];
        const REAL bpdr = a1r[q] + a3r[q], bpdi = a1i[q] + a3i[q];
        const REAL bmdr = a1r[q] - a3r[q], bmdi = a1i[q] - a3i[q];
        /* j*(b - d) with j = i */
        const REAL jr = -bmdi, ji = bmdr;
        y0r[q] = apcr + bpdr;
        y0i[q] = apci + bpdi;
        const REAL t1r = amcr - jr, t1i = amci - ji;
        y1r[q] = t1r * w1r - t1i * w1i;
        y1i[q] = t1r * w1i + t1i * w1r;
        const REAL t2r = apcr - bpdr, t2i = apci - bpdi;
        y2r[q] = t2r * w2r - t2i * w2i;
        y2i[q] = t2r * w2i + t2i * w2r;
        const REAL t3r = amcr + jr, t3i = amci + ji;
        y3r[q] = t3r * w3r - t3i * w3i;
        y3i[q] = t3r * w3i + t3i * w3r;
      }
    }
    w += (size_t)m * 6;
    REAL* t;
    t = ar, ar = br, br = t;
    t = ai, ai = bi, bi = t;
  }
  if (nt == 2) { /* last radix-2 stage, no twiddle */
    for (int q = 0; q < s; ++q) {
      const REAL a0r = ar[q], a0i = ai[q], a1r = ar[q + s], a1i = ai[q + s];
      br[q] = a0r + a1r;
      bi[q] = a0i + a1i;
      br[q + s] = a0r - a1r;
      bi[q + s] = a0i - a1i;
    }
    REAL* t;
    t = ar, ar = br, br = t;
    t = ai, ai = bi, bi = t;
  }
  if (ar != xr) {
    memcpy(xr, ar, sizeof(REAL) * (size_t)n);
    memcpy(xi, ai, sizeof(REAL) * (size_t)n);
  }
}

int fdas_cpu_fop(const float* x_c64, int n, const float* taps_c64, int t_count, int taps, int tile, float* fop,
                 int n_threads) {
  if (!x_c64 || !taps_c64 || !fop || n <= 0 || t_count <= 0 || taps <= 0 || tile < 2 * taps) return -1;
  if (tile & (tile - 1)) return -1;
  const int o = taps - 1, p = tile - o;
  const int n_tiles = (n + p - 1) / p;
  fft_plan pl;
  if (fft_plan_make(&pl, tile)) return -2;
  /* transformed tiles, split: re then im per tile */
  REAL* xt = (REAL*)malloc(sizeof(REAL) * 2 * (size_t)tile * (size_t)n_tiles);
  if (!xt) {
    fft_plan_free(&pl);
    return -2;
  }
#ifdef _OPENMP
  if (n_threads > 0) omp_set_num_threads(n_threads);
#endif
  int err = 0;
  /* forward transform of every tile: tile j holds X[j*P - O + s] (ft_conv.py:164-180) */
#pragma omp parallel
  {
    REAL* wk = (REAL*)malloc(sizeof(REAL) * 2 * (size_t)tile);
    if (!wk) {
#pragma omp atomic write
      err = -2;
    }
#pragma omp for schedule(static)
    for (int j = 0; j < n_tiles; ++j) {
      if (!wk) continue;
      REAL* tr = xt + (size_t)j * 2 * tile;
      REAL* ti = tr + tile;
      for (int s = 0; s < tile; ++s) {
        const long idx = (long)j * p - o + s;
        if (idx >= 0 && idx < n) {
          tr[s] = (REAL)x_c64[2 * idx];
          ti[s] = (REAL)x_c64[2 * idx + 1];
        } else {
          tr[s] = ti[s] = 0;
        }
      }
      fft_forward(&pl, tr, ti, wk, wk + tile);
    }
    free(wk);
  }
  if (err) {
    free(xt);
    fft_plan_free(&pl);
    return err;
  }
#pragma omp parallel
  {
    REAL* buf = (REAL*)malloc(sizeof(REAL) * 6 * (size_t)tile); /* h (split), y (split), work (split) */
    if (!buf) {
#pragma omp atomic write
      err = -2;
    }
#pragma omp for schedule(dynamic, 1)
    for (int r = 0; r < t_count; ++r) {
      if (!buf) continue;
      REAL *hr = buf, *hi = buf + tile, *yr = buf + 2 * tile, *yi = buf + 3 * tile, *wr = buf + 4 * tile,
           *wi = buf + 5 * tile;
      for (int s = 0; s < tile; ++s) {
        hr[s] = hi[s] = 0;
        if (s < taps) {
          hr[s] = (REAL)taps_c64[2 * ((size_t)r * taps + s)];
          hi[s] = (REAL)taps_c64[2 * ((size_t)r * taps + s) + 1];
        }
      }
      fft_forward(&pl, hr, hi, wr, wi);
      /* fold the 1/S of the inverse into the template spectrum */
      const REAL inv = (REAL)1 / (REAL)tile;
      for (int s = 0; s < tile; ++s) {
        hr[s] *= inv;
        hi[s] *= inv;
      }
      for (int j = 0; j < n_tiles; ++j) {
        const REAL* tr = xt + (size_t)j * 2 * tile;
        const REAL* ti = tr + tile;
        /* inverse DFT as conj(DFT(conj(.))): conjugated product in, conjugate ignored out (|y|^2) */
        for (int s = 0; s < tile; ++s) {
          yr[s] = tr[s] * hr[s] - ti[s] * hi[s];
          yi[s] = -(tr[s] * hi[s] + ti[s] * hr[s]);
        }
        fft_forward(&pl, yr, yi, wr, wi);
        /* sample s >= O of tile j is FOP bin j*P + s - O; bins past N are dropped */
        const long f0 = (long)j * p - o;
        const int s_end = f0 + tile > n ? (int)(n - f0) : tile;
        float* row = fop + (size_t)r * n;
        for (int s = o; s < s_end; ++s) row[f0 + s] = (float)(yr[s] * yr[s] + yi[s] * yi[s]);
      }
    }
    free(buf);
  }
  free(xt);
  fft_plan_free(&pl);
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
  /* Rows in parallel; each row in chunks of CHUNK bins so the partial sums stay
   * in cache while the planes are added (the own row and the sub-rows are read
   * once per chunk).  Detections of a (row, plane) are appended in bin order. */
  enum { CHUNK = 8192 };
#pragma omp parallel
  {
    REAL* hp = (REAL*)malloc(sizeof(REAL) * CHUNK);
    size_t* capl = (size_t*)calloc((size_t)n_harm, sizeof(size_t));
#pragma omp for schedule(dynamic, 1)
    for (int r = 0; r < t_count; ++r) {
      if (!hp || !capl) {
#pragma omp atomic write
        err = -2;
        continue;
      }
      const int t = r - c;
      cand** lr = lists + (size_t)r * n_harm;
      size_t* nr = lens + (size_t)r * n_harm;
      for (int k = 0; k < n_harm; ++k) capl[k] = 0;
      for (int f0 = 0; f0 < n; f0 += CHUNK) {
        const int m = n - f0 < CHUNK ? n - f0 : CHUNK;
        for (int k = 1; k <= n_harm; ++k) {
          const float* src = fop + (size_t)(c + t / k) * n; /* C division truncates toward zero */
          if (k == 1) {
            for (int i = 0; i < m; ++i) hp[i] = (REAL)src[f0 + i];
          } else {
            /* bins f .. f+k-1 with f a multiple of k share sub-harmonic bin f/k */
            int i = 0;
            for (; i < m && (f0 + i) % k; ++i) hp[i] += (REAL)src[(f0 + i) / k];
            for (; i + k <= m; i += k) {
              const REAL v = (REAL)src[(f0 + i) / k];
              for (int u = 0; u < k; ++u) hp[i + u] += v;
            }
            for (; i < m; ++i) hp[i] += (REAL)src[(f0 + i) / k];
          }
          const REAL th = (REAL)thr[k - 1];
          for (int i = 0; i < m; ++i) {
            if (!(hp[i] > th)) continue;
            size_t* cnt = &nr[k - 1];
            if (*cnt == capl[k - 1]) {
              capl[k - 1] = capl[k - 1] ? 2 * capl[k - 1] : 64;
              cand* nl = (cand*)realloc(lr[k - 1], capl[k - 1] * sizeof(cand));
              if (!nl) {
#pragma omp atomic write
                err = -2;
                break;
              }
              lr[k - 1] = nl;
            }
            lr[k - 1][*cnt].pwr = (float)hp[i];
            lr[k - 1][*cnt].key = ((uint32_t)r << 22) | (uint32_t)(f0 + i);
            ++*cnt;
          }
        }
      }
    }
    free(hp);
    free(capl);
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
