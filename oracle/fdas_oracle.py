"""CPU oracle for the FDAS hot path (test infrastructure only).

This module restates, in NumPy/SciPy and in FP64 unless stated otherwise, the
reference's own CPU path for Fourier-domain acceleration search:

  spectrum (ft_conv.py) -> overlap-save FT convolution -> filter-output plane
  (FOP) -> harmonic summing (hsum.py) -> thresholding -> candidate lists.

It is test infrastructure, never the thing measured or shipped: only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it.  The GPU product path in ``fdas_amd`` never calls into this file.

Every function cites the reference file:line it follows (paths relative to the
UOA-PARC-SKA/FDAS repository root).  How each function is pinned (golden data
held by the reference, closed forms, brute force) is listed in
``PINNING`` at the bottom of this file and exercised by ``tests/test_oracle.py``.
"""

from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np
import scipy.special

# ---------------------------------------------------------------------------
# Geometry  (FDAS.cpp:177-186, gen_info.h:41-46, cl_codegen.py:99-100,
#            ft_conv.py:159-163)
# ---------------------------------------------------------------------------


def overlap_for(max_taps: int, pack: int = 4) -> int:
    """Tile overlap O = ceil((L_max - 1) / pack) * pack  (cl_codegen.py:99, ft_conv.py:161)."""
    return int(math.ceil((max_taps - 1) / pack)) * pack


@dataclass(frozen=True)
class Geometry:
    n_bins: int
    tile: int
    overlap: int
    payload: int
    n_tiles: int
    padding_last_tile: int

    @property
    def tiles_sz(self) -> int:
        return self.n_tiles * self.tile


def geometry(n_bins: int, tile: int = 2048, max_taps: int = 421) -> Geometry:
    """n_tiles = ceil(N / P), padding of the last tile (FDAS.cpp:177-184)."""
    o = overlap_for(max_taps)
    p = tile - o
    if p <= 0:
        raise ValueError("tile too small for template length")
    n_tiles = n_bins // p + 1
    pad = p - n_bins % p
    if pad == p:
        n_tiles -= 1
        pad = 0
    return Geometry(n_bins, tile, o, p, n_tiles, pad)


# ---------------------------------------------------------------------------
# Spectrum  (ft_conv.py:134-143)
# ---------------------------------------------------------------------------


def spectrum(samples: np.ndarray, n_bins: int) -> np.ndarray:
    """X[f] = DFT(s)[f] / Ns for f < N, stored as complex64 (ft_conv.py:134-143).

    The reference resizes the series to Ns = 2N samples first (ft_conv.py:128-132).
    The transform here is taken in FP64 and rounded once to complex64.
    """
    ns = 2 * n_bins
    s = np.asarray(samples, dtype=np.float64)
    if s.size < ns:
        raise ValueError("not enough samples")
    s = s[:ns]
    ft = np.fft.fft(s) / ns
    return ft[:n_bins].astype(np.complex64)


# ---------------------------------------------------------------------------
# Template bank  (tmpl-gen.py:28-78, 138-159)
# ---------------------------------------------------------------------------


def _fresnel_any(a: np.ndarray):
    """Fresnel S, C of a purely real or purely imaginary argument.

    Follows tmpl-gen.py:28-42: for a = i*y the reference uses S(iy) = -i S(y)
    and C(iy) = i C(y) (the odd-function identities of the Fresnel integrals).
    """
    re_any = np.any(np.real(a))
    im_any = np.any(np.imag(a))
    if not re_any and im_any:
        s, c = scipy.special.fresnel(np.imag(a))
        return np.conj(s * 1j), c * 1j
    if re_any and not im_any:
        s, c = scipy.special.fresnel(np.real(a))
        return s + 0j, c + 0j
    raise RuntimeError("argument neither purely real nor purely imaginary")


def drift_response(z: float) -> np.ndarray:
    """Fourier response of a tone drifting by z bins (tmpl-gen.py:56-76).

    w = arange(-|z|, |z|) - z/2;  Y = w*sqrt(2/z);  Z = (w+z)*sqrt(2/z)
    h = (2z)^(-1/2) * exp(i*pi*w^2/z) * (S(Z) - S(Y) - i*(C(Y) - C(Z)))
    z == 0 gives the single tap 1 (tmpl-gen.py:75-76).
    """
    if z == 0:
        return np.ones(1, dtype=np.complex128)
    w = np.arange(-abs(z), abs(z)) - z / 2.0
    root = np.sqrt(complex(2.0 / z))
    y = w * root
    zz = (w + z) * root
    s_y, c_y = _fresnel_any(y)
    s_z, c_z = _fresnel_any(zz)
    return (1.0 / np.sqrt(complex(2.0 * z))) * np.exp(1j * np.pi * np.square(w) / z) \
        * (s_z - s_y - 1j * (c_y - c_z))


def reference_drifts(n_templates: int = 85, max_accel: float = 350.0, unaccel_freq: float = 500.0,
                     t_obs: float = 600.0, c_light: float = 299792458.0) -> np.ndarray:
    """Drift z_t (bins) of the reference bank generator (tmpl-gen.py:52-56, 127-133)."""
    num = n_templates - 1
    step = 2.0 * max_accel / num
    return np.array([(step * j - max_accel) * unaccel_freq * t_obs ** 2 / c_light for j in range(num + 1)])


def make_bank(drifts, max_taps: int | None = None, dtype=np.complex64) -> np.ndarray:
    """Zero-padded (T, max_taps) bank (tmpl-gen.py:154-159)."""
    resp = [drift_response(float(z)) for z in drifts]
    width = max(r.size for r in resp) if max_taps is None else max_taps
    out = np.zeros((len(resp), width), dtype=dtype)
    for i, r in enumerate(resp):
        if r.size > width:
            raise ValueError("template longer than max_taps")
        out[i, :r.size] = r
    return out


def bit_rev(x: int, bits: int) -> int:
    y = 0
    for _ in range(bits):
        y = (y << 1) | (x & 1)
        x >>= 1
    return y


def to_fft_order(spectra: np.ndarray, p: int = 4) -> np.ndarray:
    """Permute length-S rows into p-parallel FFT order (tmpl-gen.py:145-153, ft_conv.py:185-194).

    out[j::p] = in[bitrev(j)*S/p : (bitrev(j)+1)*S/p].
    """
    spectra = np.asarray(spectra)
    s = spectra.shape[-1]
    bits = int(round(math.log2(p)))
    out = np.empty_like(spectra)
    for j in range(p):
        b = bit_rev(j, bits) * s // p
        out[..., j::p] = spectra[..., b:b + s // p]
    return out


def from_fft_order(spectra: np.ndarray, p: int = 4) -> np.ndarray:
    """Inverse of :func:`to_fft_order`."""
    spectra = np.asarray(spectra)
    s = spectra.shape[-1]
    bits = int(round(math.log2(p)))
    out = np.empty_like(spectra)
    for j in range(p):
        b = bit_rev(j, bits) * s // p
        out[..., b:b + s // p] = spectra[..., j::p]
    return out


def bank_spectra(bank: np.ndarray, tile: int) -> np.ndarray:
    """DFT_S of each zero-padded template (tmpl-gen.py:139-143), FP64."""
    return np.fft.fft(np.asarray(bank, dtype=np.complex128), tile, axis=-1)


# ---------------------------------------------------------------------------
# Overlap-save tiling and the forward tile transform  (ft_conv.py:158-183)
# ---------------------------------------------------------------------------


def tiles(x: np.ndarray, geo: Geometry) -> np.ndarray:
    """Tile j holds X[j*P - O + s], zero outside [0, N) (ft_conv.py:164-180)."""
    x = np.asarray(x, dtype=np.complex128)
    out = np.zeros((geo.n_tiles, geo.tile), dtype=np.complex128)
    for j in range(geo.n_tiles):
        lo = j * geo.payload - geo.overlap
        a = max(lo, 0)
        b = min(lo + geo.tile, geo.n_bins)
        out[j, a - lo:b - lo] = x[a:b]
    return out


def tiles_transformed(x: np.ndarray, geo: Geometry) -> np.ndarray:
    """Unnormalised forward DFT of each tile, natural order (ft_conv.py:183)."""
    return np.fft.fft(tiles(x, geo), axis=-1)


# ---------------------------------------------------------------------------
# Filter-output plane  (ft_conv.py:199-207)
# ---------------------------------------------------------------------------


def part_rows(n_templates: int, part: str) -> range:
    """Rows of a FOP part (FDAS.cpp:331-346): 'neg' = t in [-c,0], 'pos' = [0,c], 'all'."""
    c = n_templates // 2
    if part == "neg":
        return range(0, c + 1)
    if part == "pos":
        return range(c, 2 * c + 1)
    if part == "all":
        return range(0, 2 * c + 1)
    raise ValueError(part)


def fop_direct(x: np.ndarray, bank: np.ndarray, rows=None) -> np.ndarray:
    """Brute force FOP[t][f] = |sum_m h_t[m] X[f-m]|^2, X[<0] = 0 (ft_conv.py:201-205), FP64.

    O(N * L) per row; intended for small N.  Rows outside ``rows`` stay 0.
    """
    x = np.asarray(x, dtype=np.complex128)
    n = x.size
    bank = np.asarray(bank, dtype=np.complex128)
    t_count = bank.shape[0]
    out = np.zeros((t_count, n), dtype=np.float64)
    rows = range(t_count) if rows is None else rows
    for r in rows:
        acc = np.zeros(n, dtype=np.complex128)
        for m, h in enumerate(bank[r]):
            if m >= n:
                break
            if h == 0:
                continue
            acc[m:] += h * x[:n - m]
        out[r] = acc.real ** 2 + acc.imag ** 2
    return out


def fop_fft(x: np.ndarray, bank: np.ndarray, rows=None, block: int = 8192) -> np.ndarray:
    """FOP via FP64 overlap-save with FFT size ``block`` (same value as fop_direct).

    Causal linear convolution truncated to N equals overlap-save whenever the
    overlap is at least L - 1; FP64 keeps the difference to direct summation
    at the 1e-15 relative level.  Used for sizes where fop_direct is too slow.
    """
    x = np.asarray(x, dtype=np.complex128)
    n = x.size
    bank = np.asarray(bank, dtype=np.complex128)
    t_count, width = bank.shape
    o = width - 1
    if block <= o:
        block = 1 << int(math.ceil(math.log2(4 * (o + 1))))
    p = block - o
    n_tiles = (n + p - 1) // p
    padded = np.zeros(o + n_tiles * p + o, dtype=np.complex128)
    padded[o:o + n] = x
    idx = np.arange(n_tiles)[:, None] * p + np.arange(block)[None, :]
    xt = np.fft.fft(padded[idx], axis=-1)
    hs = np.fft.fft(bank, block, axis=-1)
    out = np.zeros((t_count, n), dtype=np.float64)
    rows = range(t_count) if rows is None else rows
    for r in rows:
        y = np.fft.ifft(xt * hs[r][None, :], axis=-1)[:, o:]
        y = y.reshape(-1)[:n]
        out[r] = y.real ** 2 + y.imag ** 2
    return out


# ---------------------------------------------------------------------------
# Harmonic summing  (hsum.py:91-120, hms_baseline.cl.mako:55-71)
# ---------------------------------------------------------------------------


def sub_row(r: int, k: int, c: int) -> int:
    """Row of FOP[trunc(t/k)] for row r = t + c; truncation toward zero (hsum.py:109-114)."""
    t = r - c
    q = abs(t) // k
    return c + (q if t >= 0 else -q)


def harmonic_planes(fop: np.ndarray, n_harmonics: int, rows=None, dtype=np.float64) -> np.ndarray:
    """HP_1 = FOP; HP_k = HP_{k-1} + FOP[trunc(t/k)][f // k], summed in order (hsum.py:93-120).

    ``dtype`` selects the arithmetic (float32 restates the reference's float32
    NumPy additions exactly; float64 is the FP64 oracle).  Rows outside
    ``rows`` are left at 0.  Returns an (H, T, N) array.
    """
    fop = np.asarray(fop, dtype=dtype)
    t_count, n = fop.shape
    c = t_count // 2
    rows = range(t_count) if rows is None else rows
    hps = np.zeros((n_harmonics,) + fop.shape, dtype=dtype)
    for r in rows:
        hps[0, r] = fop[r]
    for k in range(2, n_harmonics + 1):
        m = (n + k - 1) // k
        for r in rows:
            src = fop[sub_row(r, k, c), :m]
            hps[k - 1, r] = hps[k - 2, r] + np.repeat(src, k)[:n]
    return hps


def harmonic_planes_loop(fop: np.ndarray, n_harmonics: int) -> np.ndarray:
    """Literal per-element restatement of hsum.py:93-120 (float32, tiny inputs only).

    Keeps the reference's ceil / floor branches on the sign of the template
    number instead of the closed-form truncation used by harmonic_planes.
    """
    fop = np.asarray(fop, dtype=np.float32)
    t_count, n = fop.shape
    c = t_count // 2
    hps = np.empty((n_harmonics,) + fop.shape, dtype=np.float32)
    hps[0] = fop
    for h in range(1, n_harmonics):
        k = h + 1
        sp = None
        for tmpl in range(t_count):
            num = tmpl - c
            if num < 0:
                sp = int(np.ceil(num / k)) + c
            elif num > 0:
                sp = int(np.floor(num / k)) + c
            else:  # the reference keeps the previous value here, which is c (hsum.py:113-114)
                pass
            for freq in range(n):
                hps[h, tmpl, freq] = np.float32(hps[h - 1, tmpl, freq] + fop[sp, freq // k])
    return hps


# ---------------------------------------------------------------------------
# Thresholds  (hsum.py:54-73, 125-131)
# ---------------------------------------------------------------------------


def noise_thresholds(x: np.ndarray, n_harmonics: int, sigma: float = 6.0) -> np.ndarray:
    """T_k = (k + sigma*sqrt(k)) * mean|X|^2, mean taken in FP64, stored float32."""
    x = np.asarray(x, dtype=np.complex128)
    p = float(np.mean(x.real ** 2 + x.imag ** 2))
    k = np.arange(1, n_harmonics + 1, dtype=np.float64)
    return ((k + sigma * np.sqrt(k)) * p).astype(np.float32)


def quantile_thresholds(hps: np.ndarray, n_cands_per_plane: int, rows=None) -> np.ndarray:
    """Mock thresholds thr_k = quantile(HP_k, 1 - n/M) (hsum.py:125-131), M values per plane.

    Linear interpolation at q (M - 1) between order statistics, evaluated in FP64
    and rounded once to float32 (what the reference's NumPy 1.19 computes for a
    float32 plane; NumPy 2 would round the index to float32 first).  ``rows``
    restricts the planes to a FOP part.
    """
    hps = np.asarray(hps)
    if rows is not None:
        hps = hps[:, list(rows)]
    size = hps[0].size
    return np.array([np.quantile(hps[h].astype(np.float64), 1.0 - n_cands_per_plane / size)
                     for h in range(hps.shape[0])], dtype=np.float32)


# ---------------------------------------------------------------------------
# Detection, selection and encoding  (hsum.py:133-151, gen_info.h:62-73,
#                                     detect.cl.mako:81-143)
# ---------------------------------------------------------------------------

INVALID_POWER = np.float32(-1.0)


def encode_location(k, t, f, c: int = 42):
    """((k-1)&7)<<29 | ((t+c)&0x7f)<<22 | (f&0x3fffff)  (gen_info.h:62-63, hsum.py:147)."""
    k = np.asarray(k, dtype=np.int64)
    t = np.asarray(t, dtype=np.int64)
    f = np.asarray(f, dtype=np.int64)
    v = (((k - 1) & 0x7) << 29) | (((t + c) & 0x7F) << 22) | (f & 0x3FFFFF)
    return v.astype(np.uint32)


def decode_location(loc, c: int = 42):
    """(harmonic, template number, bin) of a 32-bit location (gen_info.h:65-67)."""
    loc = np.asarray(loc, dtype=np.uint64)
    k = ((loc >> 29) & 0x7) + 1
    t = ((loc >> 22) & 0x7F).astype(np.int64) - c
    f = loc & 0x3FFFFF
    return k.astype(np.int64), t, f.astype(np.int64)


def invalid_location(c: int = 42) -> np.uint32:
    """encode(1, c+1, 0): 0x15400000 for c = 42 (gen_info.h:69)."""
    return np.uint32(encode_location(1, c + 1, 0, c))


@dataclass
class PlaneCandidates:
    """Candidates of one harmonic plane: rows (t + c), bins, powers; sorted by (row, bin)."""
    rows: np.ndarray
    bins: np.ndarray
    powers: np.ndarray
    n_detected: int


def detect(hps: np.ndarray, thresholds, rows=None) -> list[PlaneCandidates]:
    """All (row, bin) with HP_k > thr_k, strict (hsum.py:133-151, detect.cl.mako:81-83).

    Listed plane-major, then row-major - the reference's own output order.
    """
    h_count, t_count, n = hps.shape
    rows = np.arange(t_count) if rows is None else np.asarray(list(rows))
    out = []
    for h in range(h_count):
        sub = hps[h][rows]
        rr, ff = np.nonzero(sub > np.asarray(thresholds[h], dtype=sub.dtype))
        rr = rows[rr]
        out.append(PlaneCandidates(rr.astype(np.int64), ff.astype(np.int64),
                                   hps[h][rr, ff], int(rr.size)))
    return out


def select(cands: list[PlaneCandidates], cap: int) -> list[PlaneCandidates]:
    """Cap each plane at ``cap`` entries: strongest power first, ties by (row, bin).

    The reference FPGA keeps a 64-slot ring per plane and silently overwrites
    (detect.cl.mako:106-115); the library replaces that with a deterministic
    rule so that GPU and oracle lists can be compared entry for entry.  The
    kept entries are returned sorted by (row, bin).
    """
    out = []
    for pc in cands:
        if pc.rows.size > cap:
            order = np.lexsort((pc.bins, pc.rows, -pc.powers.astype(np.float64)))[:cap]
            rows, bins, pw = pc.rows[order], pc.bins[order], pc.powers[order]
        else:
            rows, bins, pw = pc.rows, pc.bins, pc.powers
        o = np.lexsort((bins, rows))
        out.append(PlaneCandidates(rows[o], bins[o], pw[o], pc.n_detected))
    return out


def reference_lists(sel: list[PlaneCandidates], cap: int, c: int = 42):
    """(loc, pwr) arrays in the reference's plane-major layout with invalid padding.

    Plane h owns entries [h*cap, (h+1)*cap); unused entries are encode(1,c+1,0)
    and -1.0f (gen_info.h:69, detect.cl.mako:38-39, store_cands.cl.mako:19-33).
    """
    h_count = len(sel)
    loc = np.full(h_count * cap, invalid_location(c), dtype=np.uint32)
    pwr = np.full(h_count * cap, INVALID_POWER, dtype=np.float32)
    for h, pc in enumerate(sel):
        m = min(cap, pc.rows.size)
        loc[h * cap:h * cap + m] = encode_location(h + 1, pc.rows[:m] - c, pc.bins[:m], c)
        pwr[h * cap:h * cap + m] = pc.powers[:m]
    return loc, pwr


# ---------------------------------------------------------------------------
# Whole path
# ---------------------------------------------------------------------------


def run(x: np.ndarray, bank: np.ndarray, thresholds, n_harmonics: int, cap: int, part: str = "all",
        fop=None, method: str = "auto"):
    """FOP -> planes -> detect -> select for one FOP part, all in FP64."""
    t_count = bank.shape[0]
    rows = part_rows(t_count, part)
    if fop is None:
        if method == "direct" or (method == "auto" and x.size * bank.shape[1] <= 4_000_000):
            fop = fop_direct(x, bank, rows)
        else:
            fop = fop_fft(x, bank, rows)
    hps = harmonic_planes(fop, n_harmonics, rows)
    sel = select(detect(hps, thresholds, rows), cap)
    return fop, hps, sel


PINNING = {
    "geometry": "closed forms; table values 21/1289/2577 tiles and paddings 1420/1340/1052 (tests)",
    "spectrum": "pure tone at an integer bin -> single bin of amplitude A/2; Parseval",
    "drift_response/make_bank": "reference fixture fdas_templates_85_350.npy (85x421)",
    "to_fft_order/bank_spectra": "reference fixture fdas_templates_85_350_ft_p4.npy (85x2048)",
    "tiles": "ft_conv.py:164-180 geometry: overlap copy of previous tile, zero head/tail",
    "fop_direct": "t=0 tap 1 -> |X|^2; delta input -> |h|^2 shifted; linearity",
    "fop_fft": "equals fop_direct on ragged N (N < P, N = P, N not multiple of P)",
    "harmonic_planes": "literal hsum.py loop restatement (harmonic_planes_loop); constant FOP -> k*FOP",
    "quantile_thresholds": "np.quantile count property: ~n candidates per plane",
    "encode_location": "0x15400000 = encode(1, 43, 0) from gen_info.h:69; decode(encode) = id",
    "select": "brute force top-k by sorting all candidates",
}


def detect_streaming(fop: np.ndarray, n_harmonics: int, thresholds, rows=None, keep_planes=()):
    """detect(harmonic_planes(...)) holding one FP64 plane at a time (for N = 2^21 and up).

    Returns (candidates per plane, {k: plane} for k in keep_planes).
    """
    fop = np.asarray(fop, dtype=np.float64)
    t_count, n = fop.shape
    c = t_count // 2
    rows = np.arange(t_count) if rows is None else np.asarray(list(rows))
    hp = np.zeros_like(fop)
    hp[rows] = fop[rows]
    out, kept = [], {}
    for k in range(1, n_harmonics + 1):
        if k > 1:
            m = (n + k - 1) // k
            for r in rows:
                hp[r] += np.repeat(fop[sub_row(int(r), k, c), :m], k)[:n]
        if k in keep_planes:
            kept[k] = hp.copy()
        sub = hp[rows]
        rr, ff = np.nonzero(sub > np.float64(thresholds[k - 1]))
        rr = rows[rr]
        out.append(PlaneCandidates(rr.astype(np.int64), ff.astype(np.int64), hp[rr, ff], int(rr.size)))
    return out, kept
