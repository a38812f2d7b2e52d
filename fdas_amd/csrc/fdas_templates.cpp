// Template-bank generation and host-side transforms (FP64).
//
// The accelerated-response templates follow the reference generator
// fdas/tmpl/tmpl-gen.py:45-78 (itself a port of the SKA TDT Matlab model).
// The Fresnel integrals are this library's own, from their mathematics:
//   * |x| < 1.5 : the power series C(x) = sum_n (-1)^n (pi/2)^(2n) x^(4n+1) / ((2n)! (4n+1)),
//                 S(x) = sum_n (-1)^n (pi/2)^(2n+1) x^(4n+3) / ((2n+1)! (4n+3));
//   * |x| >= 1.5: C + iS = (1+i)/2 * erf(w), w = (sqrt(pi)/2)(1-i)x, with erfc(w)
//                 from the even contraction of Laplace's continued fraction
//                 erfc(w) = exp(-w^2)/sqrt(pi) * 2w / (2w^2+1 - 1*2/(2w^2+5 - 3*4/(2w^2+9 - ...))),
//                 evaluated with the modified Lentz algorithm (Lentz 1976;
//                 Thompson & Barnett 1986).
#include <cmath>
#include <complex>
#include <cstdint>
#include <vector>

#include "fdas_host.h"

namespace fdas {

namespace {

constexpr double kPi = 3.14159265358979323846264338327950288;

void fresnel_series(double x, double* s, double* c) {
  const double t = 0.5 * kPi * x * x;
  double a = x;  // a_k = (pi/2)^k x^(2k+1) / k!
  double cs = 0.0, ss = 0.0;
  for (int k = 0; k < 200; ++k) {
    const double term = a / double(2 * k + 1);
    switch (k & 3) {
      case 0: cs += term; break;
      case 1: ss += term; break;
      case 2: cs -= term; break;
      default: ss -= term; break;
    }
    if (k > 2 && std::fabs(term) < 1e-18 * (std::fabs(cs) + std::fabs(ss))) break;
    a *= t / double(k + 1);
  }
  *s = ss;
  *c = cs;
}

void fresnel_cf(double x, double* s, double* c) {
  using cd = std::complex<double>;
  const double tiny = 1e-300;
  const cd b1(1.0, -kPi * x * x);  // 2w^2 + 1 with 2w^2 = -i*pi*x^2
  cd f = b1, C = b1, D = 0.0;
  for (int n = 2; n < 100000; ++n) {
    const double an = -double(2 * n - 3) * double(2 * n - 2);
    const cd bn = b1 + 4.0 * double(n - 1);
    D = bn + an * D;
    if (std::abs(D) < tiny) D = tiny;
    D = 1.0 / D;
    C = bn + an / C;
    if (std::abs(C) < tiny) C = tiny;
    const cd delta = C * D;
    f *= delta;
    if (std::abs(delta - 1.0) < 1e-16) break;
  }
  // erfc(w) = exp(i*pi*x^2/2) * x(1-i) / f
  const double th = 0.5 * kPi * x * x;
  const cd erfc_w = std::polar(1.0, th) * cd(x, -x) / f;
  const cd cs = cd(0.5, 0.5) * (1.0 - erfc_w);
  *c = cs.real();
  *s = cs.imag();
}

}  // namespace

void fresnel(double x, double* s, double* c) {
  const double ax = std::fabs(x);
  if (ax < 1.5) fresnel_series(ax, s, c);
  else fresnel_cf(ax, s, c);
  if (x < 0) {
    *s = -*s;
    *c = -*c;
  }
}

// Number of taps of numpy.arange(-|z|, |z|) (tmpl-gen.py:61).
uint32_t drift_taps(double z) {
  if (z == 0.0) return 1;
  const double az = std::fabs(z);
  return uint32_t(std::ceil((az - (-az)) / 1.0));
}

// Fourier response of a tone drifting by z bins (tmpl-gen.py:56-76).
std::vector<std::complex<double>> drift_response(double z) {
  using cd = std::complex<double>;
  if (z == 0.0) return {cd(1.0, 0.0)};
  const double az = std::fabs(z);
  const uint32_t n = drift_taps(z);
  // numpy.arange fills start + i * (x[1] - x[0]).
  const double start = -az;
  const double delta = (start + 1.0) - start;
  std::vector<cd> h(n);
  const double sq = std::sqrt(2.0 / az);
  const double inv = 1.0 / std::sqrt(2.0 * az);
  for (uint32_t m = 0; m < n; ++m) {
    const double w = (start + double(m) * delta) - z / 2.0;
    const double y = w * sq, zz = (w + z) * sq;
    double sy, cy, sz, cz;
    fresnel(y, &sy, &cy);
    fresnel(zz, &sz, &cz);
    const cd ph = std::polar(1.0, kPi * w * w / z);
    cd val;
    if (z > 0) {
      // (2z)^(-1/2) * (S(Z) - S(Y) - i (C(Y) - C(Z)))
      val = inv * ph * cd(sz - sy, -(cy - cz));
    } else {
      // Y, Z purely imaginary: S(iy) = -i S(y), C(iy) = i C(y); (2z)^(-1/2) = -i / sqrt(2|z|)
      const cd SY(0.0, -sy), SZ(0.0, -sz), CY(0.0, cy), CZ(0.0, cz);
      val = cd(0.0, -inv) * ph * (SZ - SY - cd(0.0, 1.0) * (CY - CZ));
    }
    h[m] = val;
  }
  return h;
}

void fft_inplace(std::complex<double>* a, uint32_t n, int dir) {
  // iterative radix-2, n a power of two; dir = -1 forward, +1 inverse (unnormalised)
  for (uint32_t i = 1, j = 0; i < n; ++i) {
    uint32_t bit = n >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j) std::swap(a[i], a[j]);
  }
  for (uint32_t len = 2; len <= n; len <<= 1) {
    const double ang = dir * 2.0 * kPi / double(len);
    for (uint32_t i = 0; i < n; i += len) {
      for (uint32_t k = 0; k < len / 2; ++k) {
        const std::complex<double> w = std::polar(1.0, ang * double(k));
        const std::complex<double> u = a[i + k], v = a[i + k + len / 2] * w;
        a[i + k] = u + v;
        a[i + k + len / 2] = u - v;
      }
    }
  }
}

}  // namespace fdas

extern "C" fdas_status fdas_make_drift_bank(const double* drifts, uint32_t n, uint32_t max_taps, float* out_c64) {
  if (!drifts || !out_c64 || n == 0 || max_taps == 0) return FDAS_EINVAL;
  for (uint32_t i = 0; i < n; ++i)
    if (!std::isfinite(drifts[i]) || fdas::drift_taps(drifts[i]) > max_taps) return FDAS_EINVAL;
  for (uint32_t i = 0; i < n; ++i) {
    const auto h = fdas::drift_response(drifts[i]);
    float* row = out_c64 + size_t(i) * max_taps * 2;
    for (uint32_t m = 0; m < max_taps; ++m) {
      const std::complex<double> v = m < h.size() ? h[m] : std::complex<double>(0.0, 0.0);
      row[2 * m] = float(v.real());
      row[2 * m + 1] = float(v.imag());
    }
  }
  return FDAS_OK;
}

extern "C" fdas_status fdas_reference_drifts(uint32_t n_templates, double max_accel, double unaccel_freq,
                                             double t_obs, double speed_of_light, double* out) {
  if (!out || n_templates < 2 || speed_of_light == 0.0) return FDAS_EINVAL;
  const uint32_t num = n_templates - 1;
  const double step = 2.0 * max_accel / double(num);
  for (uint32_t j = 0; j <= num; ++j)
    out[j] = (step * double(j) - max_accel) * unaccel_freq * (t_obs * t_obs) / speed_of_light;
  return FDAS_OK;
}
