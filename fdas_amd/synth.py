"""Synthetic dedispersed time series, spectra and banks for tests and benchmarks.

Counter-based noise (SplitMix64 -> Box-Muller in FP64 -> float32) so that any
sample of any seed can be regenerated independently, plus drifting tones

    s[n] += A * sum_{h=1..Hs} cos(2*pi*h*(f0*n/Ns + z*n^2/(2*Ns^2)))

The spectrum handed to the search follows the reference's front end
(ft_conv.py:134-143): X = DFT(s)[:N] / Ns with Ns = 2N, as complex64.
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, start: int, count: int) -> np.ndarray:
    """Outputs start .. start+count-1 of the SplitMix64 stream of ``seed``."""
    with np.errstate(over="ignore"):
        i = np.arange(start + 1, start + count + 1, dtype=np.uint64)
        z = np.uint64(seed & 0xFFFFFFFFFFFFFFFF) + i * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        return z ^ (z >> np.uint64(31))


def normal_noise(seed: int, count: int) -> np.ndarray:
    """N(0,1) float32 samples, Box-Muller over consecutive SplitMix64 pairs."""
    pairs = (count + 1) // 2
    u = (splitmix64(seed, 0, 2 * pairs) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    u1, u2 = u[0::2], u[1::2]
    r = np.sqrt(-2.0 * np.log1p(-u1))
    th = 2.0 * np.pi * u2
    out = np.empty(2 * pairs, dtype=np.float64)
    out[0::2] = r * np.cos(th)
    out[1::2] = r * np.sin(th)
    return out[:count].astype(np.float32)


@dataclass
class Tone:
    f0: float      # fundamental bin at n = 0
    z: float       # drift in bins over the series
    harmonics: int
    amplitude: float


def time_series(n_samples: int, seed: int, tones=()) -> np.ndarray:
    s = normal_noise(seed, n_samples).astype(np.float64)
    n = np.arange(n_samples, dtype=np.float64)
    ns = float(n_samples)
    for tn in tones:
        phase = tn.f0 * n / ns + tn.z * n * n / (2.0 * ns * ns)
        for h in range(1, tn.harmonics + 1):
            s += tn.amplitude * np.cos(2.0 * np.pi * h * phase)
    return s.astype(np.float32)


def spectrum(series: np.ndarray, n_bins: int) -> np.ndarray:
    """X[f] = DFT(s)[f] / Ns, f < N, Ns = 2N (ft_conv.py:134-143), complex64."""
    ns = 2 * n_bins
    s = np.asarray(series[:ns], dtype=np.float64)
    return (np.fft.rfft(s)[:n_bins] / ns).astype(np.complex64)


def drift_grid(n_templates: int, step: float = 2.0) -> np.ndarray:
    """z_j = step * (j - c), j = 0..T-1 (85 templates: [-84, 84]; 169: [-168, 168])."""
    c = n_templates // 2
    return step * (np.arange(n_templates, dtype=np.float64) - c)


def noise_thresholds(x: np.ndarray, n_harmonics: int, sigma: float = 6.0) -> np.ndarray:
    """T_k = (k + sigma*sqrt(k)) * mean|X|^2 in absolute power units, float32."""
    x = np.asarray(x)
    p = float(np.mean(x.real.astype(np.float64) ** 2 + x.imag.astype(np.float64) ** 2))
    k = np.arange(1, n_harmonics + 1, dtype=np.float64)
    return ((k + sigma * np.sqrt(k)) * p).astype(np.float32)


@dataclass
class Workload:
    """One search request: spectrum parameters plus search geometry."""
    name: str
    n_bins: int
    seed: int
    tones: list = field(default_factory=list)
    n_templates: int = 85
    max_taps: int = 421
    tile: int = 2048
    n_harmonics: int = 8
    cands_per_harmonic: int = 64

    def series(self) -> np.ndarray:
        return time_series(2 * self.n_bins, self.seed, self.tones)

    def spectrum(self) -> np.ndarray:
        return spectrum(self.series(), self.n_bins)


def emulator_workload() -> Workload:
    """Emulator-size search: 65,536 samples -> 32,768 bins, tone at bin 5,000, z=+20, 4 harmonics."""
    return Workload("emulator N=32768 T=85 H=8", 32768, 1, [Tone(5000.0, 20.0, 4, 0.05)],
                    cands_per_harmonic=64)


def full_workload() -> Workload:
    """Full single-trial search: N=2^21 bins, tone at bin 300,000, z=-36, 8 harmonics."""
    return Workload("single trial N=2^21 T=85 H=8", 1 << 21, 2, [Tone(300000.0, -36.0, 8, 0.02)],
                    cands_per_harmonic=4096)


def random_trial(seed: int, n_bins: int, rng_tones: int | None = None, z_max: float = 84.0) -> Workload:
    """Trial with 0-2 drifting tones at random bins in [1e4, min(2e6, N)) and z in [-z_max, z_max]."""
    r = np.random.default_rng(seed)
    k = int(r.integers(0, 3)) if rng_tones is None else rng_tones
    hi = min(2e6, 0.9 * n_bins)
    tones = [Tone(float(r.uniform(1e4, hi)), float(r.uniform(-z_max, z_max)), 8, float(r.uniform(0.01, 0.03)))
             for _ in range(k)]
    return Workload(f"trial seed={seed}", n_bins, seed, tones)


def _lshr(z, k):
    """Logical right shift of int64 tensors holding uint64 bit patterns."""
    return (z >> k) & ((1 << (64 - k)) - 1)


def _as_i64(v: int) -> int:
    v &= 0xFFFFFFFFFFFFFFFF
    return v - (1 << 64) if v >= (1 << 63) else v


def device_series(w: Workload, device="cuda"):
    """``Workload.series`` computed with torch on ``device``: the same SplitMix64
    stream (uint64 arithmetic on int64 bit patterns) and Box-Muller in FP64, so
    the noise equals the host's up to the last bit of the FP64 transcendental
    functions.  Bench and test input plumbing; the search never calls it."""
    import torch
    ns = 2 * w.n_bins
    pairs = (ns + 1) // 2
    i = torch.arange(1, 2 * pairs + 1, dtype=torch.int64, device=device)
    z = _as_i64(w.seed) + i * _as_i64(int(_GOLDEN))
    z = (z ^ _lshr(z, 30)) * _as_i64(int(_M1))
    z = (z ^ _lshr(z, 27)) * _as_i64(int(_M2))
    z = z ^ _lshr(z, 31)
    u = _lshr(z, 11).to(torch.float64) * (1.0 / 9007199254740992.0)
    u1, u2 = u[0::2], u[1::2]
    r = torch.sqrt(-2.0 * torch.log1p(-u1))
    th = 2.0 * np.pi * u2
    s = torch.stack((r * torch.cos(th), r * torch.sin(th)), dim=1).reshape(-1)[:ns]
    s = s.to(torch.float32).to(torch.float64)
    n = torch.arange(ns, dtype=torch.float64, device=device)
    for tn in w.tones:
        phase = tn.f0 * n / ns + tn.z * n * n / (2.0 * ns * ns)
        for h in range(1, tn.harmonics + 1):
            s += tn.amplitude * torch.cos(2.0 * np.pi * h * phase)
    return s.to(torch.float32)


def device_spectrum(w: Workload, device="cuda"):
    """``Workload.spectrum`` on ``device`` (complex64 tensor of N bins)."""
    import torch
    s = device_series(w, device).to(torch.float64)
    return (torch.fft.rfft(s)[: w.n_bins] / (2 * w.n_bins)).to(torch.complex64).contiguous()
