"""Pins the CPU oracle against the reference's own data and against the mathematics.

Each check is chosen so that a plausible slip in the oracle (a dropped term,
a wrong sign or index, a transposed operand) fails it; none re-types the
oracle's formula.
"""

import numpy as np
import pytest

from oracle import fdas_oracle as O


# -- geometry (FDAS.cpp:177-186) -------------------------------------------

@pytest.mark.parametrize("n,tiles,pad", [(32768, 21, 1420), (1 << 19, 323, 1556), (1 << 20, 645, 1484),
                                         (1 << 21, 1289, 1340), (1 << 22, 2577, 1052), (1628, 1, 0),
                                         (3256, 2, 0), (1, 1, 1627)])
def test_geometry_table(n, tiles, pad):
    g = O.geometry(n)
    assert (g.overlap, g.payload) == (420, 1628)
    assert (g.n_tiles, g.padding_last_tile) == (tiles, pad)
    assert g.n_tiles * g.payload - g.padding_last_tile == n


# -- spectrum (ft_conv.py:134-143) ------------------------------------------

def test_spectrum_tone_and_parseval():
    n = 4096
    ns = 2 * n
    t = np.arange(ns)
    s = 0.7 * np.cos(2 * np.pi * 123 * t / ns)
    x = O.spectrum(s, n)
    assert abs(x[123] - 0.35) < 1e-6            # a cosine of amplitude A puts A/2 in its bin
    x[123] = 0
    assert np.abs(x).max() < 1e-6
    rng = np.random.default_rng(3)
    s = rng.normal(size=ns)
    full = np.fft.fft(s) / ns
    assert np.isclose(np.sum(np.abs(full) ** 2), np.mean(s ** 2))  # Parseval of the normalised DFT
    assert np.allclose(O.spectrum(s, n), full[:n].astype(np.complex64))


# -- template bank (tmpl-gen.py) ---------------------------------------------

def test_bank_equals_reference_fixture(reference_bank):
    bank = O.make_bank(O.reference_drifts(), 421)
    assert bank.shape == reference_bank.shape == (85, 421)
    assert np.array_equal(bank, reference_bank)
    taps = [int(np.nonzero(r)[0].max()) + 1 for r in reference_bank]
    assert taps[42] == 1 and reference_bank[42, 0] == 1   # t = 0: single tap 1 + 0j
    assert taps[0] == taps[84] == 421


def test_fft_order_equals_reference_fixture(reference_bank, reference_bank_ft_p4):
    spec = O.bank_spectra(reference_bank.astype(np.complex128), 2048)
    p4 = O.to_fft_order(spec)
    assert np.abs(p4 - reference_bank_ft_p4).max() < 1e-6
    assert np.allclose(O.from_fft_order(p4), spec)
    # FFT order of a 4-parallel engine: element 4m+j comes from chunk bitrev2(j)
    v = np.arange(16)
    assert list(O.to_fft_order(v)) == [0, 8, 4, 12, 1, 9, 5, 13, 2, 10, 6, 14, 3, 11, 7, 15]


def test_drift_response_energy():
    # |h|^2 sums to ~1: the template spreads a unit-power tone over the drift
    for z in (-84.0, -10.0, 2.0, 37.0, 168.0):
        h = O.drift_response(z)
        assert h.size == int(np.ceil(2 * abs(z)))
        assert 0.95 < np.sum(np.abs(h) ** 2) <= 1.0001


# -- tiling (ft_conv.py:158-183) ---------------------------------------------

@pytest.mark.parametrize("n", [1000, 1628, 5000, 6512])
def test_tiles_overlap_structure(n):
    rng = np.random.default_rng(n)
    x = (rng.normal(size=n) + 1j * rng.normal(size=n)).astype(np.complex64)
    g = O.geometry(n)
    t = O.tiles(x, g)
    assert np.all(t[0, :g.overlap] == 0)
    for j in range(1, g.n_tiles):
        assert np.array_equal(t[j, :g.overlap], t[j - 1, -g.overlap:])
    payload = t[:, g.overlap:].reshape(-1)
    assert np.array_equal(payload[:n], x.astype(np.complex128))
    assert np.all(payload[n:] == 0)
    assert np.allclose(O.tiles_transformed(x, g), np.fft.fft(t, axis=-1))


# -- FOP (ft_conv.py:199-207) -------------------------------------------------

def _rand_spectrum(n, seed):
    rng = np.random.default_rng(seed)
    return (rng.normal(size=n) + 1j * rng.normal(size=n)).astype(np.complex64)


def test_fop_direct_matches_numpy_convolve():
    x = _rand_spectrum(700, 1)
    bank = O.make_bank(np.array([-30.0, -4.0, 0.0, 7.5, 30.0]), 61)
    fop = O.fop_direct(x, bank)
    for r in range(bank.shape[0]):
        h = bank[r][np.nonzero(bank[r])]
        conv = np.convolve(x.astype(np.complex128), h.astype(np.complex128))[:x.size]
        assert np.allclose(fop[r], np.abs(conv) ** 2, rtol=1e-12, atol=1e-14)


def test_fop_special_cases(reference_bank):
    x = _rand_spectrum(900, 2)
    fop = O.fop_direct(x, reference_bank)
    assert np.allclose(fop[42], np.abs(x.astype(np.complex128)) ** 2)  # t = 0 tap is 1
    # delta input at bin f0: row t is |h_t[f - f0]|^2
    d = np.zeros(900, dtype=np.complex64)
    d[100] = 1.0
    fd = O.fop_direct(d, reference_bank)
    for r in (0, 10, 50, 84):
        h = reference_bank[r].astype(np.complex128)
        expect = np.zeros(900)
        expect[100:100 + h.size] = np.abs(h[:800]) ** 2
        assert np.allclose(fd[r], expect)
    # quadratic in the input
    assert np.allclose(O.fop_direct(3.0 * x, reference_bank), 9.0 * fop)


@pytest.mark.parametrize("n", [1, 419, 1627, 1628, 1629, 3256, 5000])
def test_fop_fft_equals_direct(n, reference_bank):
    x = _rand_spectrum(n, n)
    a = O.fop_direct(x, reference_bank)
    b = O.fop_fft(x, reference_bank)
    assert np.abs(a - b).max() <= 1e-12 * max(a.max(), 1e-30)


# -- harmonic summing (hsum.py:91-120) ---------------------------------------

def test_planes_equal_literal_reference_loop():
    with np.load(O.__file__.replace("oracle/fdas_oracle.py", "tests/golden/golden_hsum_tiny.npz")) as z:
        fop, hps = z["fop"], z["hps"]
    assert np.array_equal(O.harmonic_planes(fop, 8, dtype=np.float32), hps)
    rng = np.random.default_rng(4)
    f2 = rng.exponential(size=(11, 37)).astype(np.float32)
    assert np.array_equal(O.harmonic_planes(f2, 8, dtype=np.float32), O.harmonic_planes_loop(f2, 8))


def test_planes_closed_forms():
    const = np.full((85, 64), 3.0)
    hps = O.harmonic_planes(const, 8)
    for k in range(1, 9):
        assert np.all(hps[k - 1] == 3.0 * k)
    rng = np.random.default_rng(5)
    fop = rng.exponential(size=(85, 200))
    hps = O.harmonic_planes(fop, 16)
    assert np.array_equal(hps[0], fop)
    # row c (t = 0) only ever adds row c: HP_k[c][f] = sum_j FOP[c][f // j]
    f = np.arange(200)
    assert np.allclose(hps[15][42], sum(fop[42][f // j] for j in range(1, 17)))
    # t = -42: trunc(-42/5) = -8 -> row 34; t = 41: trunc(41/4) = 10 -> row 52
    assert O.sub_row(0, 5, 42) == 34 and O.sub_row(83, 4, 42) == 52 and O.sub_row(41, 2, 42) == 42
    assert np.allclose(hps[4][0] - hps[3][0], fop[34][f // 5])


# -- thresholds, detection, encoding (hsum.py:125-151, gen_info.h:62-69) -----

def test_encoding_known_answers():
    assert int(O.invalid_location()) == 0x15400000
    assert int(O.encode_location(8, 42, (1 << 22) - 1)) == 0xF53FFFFF  # 7<<29 | 84<<22 | 0x3fffff
    assert int(O.encode_location(1, -42, 0)) == 0
    rng = np.random.default_rng(6)
    k = rng.integers(1, 9, 1000)
    t = rng.integers(-42, 43, 1000)
    f = rng.integers(0, 1 << 22, 1000)
    kk, tt, ff = O.decode_location(O.encode_location(k, t, f))
    assert np.array_equal(kk, k) and np.array_equal(tt, t) and np.array_equal(ff, f)


def test_quantile_thresholds_give_requested_count():
    rng = np.random.default_rng(8)
    hps = O.harmonic_planes(rng.exponential(size=(85, 4096)), 8)
    thr = O.quantile_thresholds(hps, 64)
    for h in range(8):
        assert 55 <= int(np.sum(hps[h] > thr[h])) <= 72


def test_select_is_top_k_by_power_then_key():
    rng = np.random.default_rng(9)
    rows = rng.integers(0, 85, 500)
    bins = rng.integers(0, 10000, 500)
    pw = rng.choice(np.arange(50, dtype=np.float64), 500)  # many ties
    keys = list({(int(r), int(b)): p for r, b, p in zip(rows, bins, pw)}.items())
    r = np.array([k[0][0] for k in keys])
    b = np.array([k[0][1] for k in keys])
    p = np.array([k[1] for k in keys])
    sel = O.select([O.PlaneCandidates(r, b, p, r.size)], 40)[0]
    brute = sorted(keys, key=lambda kv: (-kv[1], kv[0][0], kv[0][1]))[:40]
    assert sorted((int(x), int(y)) for x, y in zip(sel.rows, sel.bins)) == sorted(k for k, _ in brute)
    order = list(zip(sel.rows.tolist(), sel.bins.tolist()))
    assert order == sorted(order)


def test_detect_is_strict():
    hps = np.array([[[1.0, 2.0, 3.0]]], dtype=np.float32)
    c = O.detect(hps, np.array([2.0], dtype=np.float32))[0]
    assert list(c.bins) == [2]


def test_streaming_detect_equals_planes():
    rng = np.random.default_rng(10)
    fop = rng.exponential(size=(85, 500))
    thr = np.array([5.0 + 2 * k for k in range(8)], dtype=np.float32)
    a = O.detect(O.harmonic_planes(fop, 8), thr)
    b, kept = O.detect_streaming(fop, 8, thr, keep_planes=(3,))
    for x, y in zip(a, b):
        assert np.array_equal(x.rows, y.rows) and np.array_equal(x.bins, y.bins)
    assert np.allclose(kept[3], O.harmonic_planes(fop, 8)[2])


def test_golden_small_is_consistent(golden_small):
    g = golden_small
    assert g["fop"].shape == (85, 3000)
    assert np.allclose(g["fop"][42], np.abs(g["spectrum"].astype(np.complex128)) ** 2, rtol=1e-6)
    for h in range(8):
        assert np.all(g[f"powers_{h + 1}"] > g["thresholds"][h])


def test_quantile_thresholds_definition():
    """Linear interpolation between order statistics at q (M - 1), FP64, one rounding to float32."""
    rng = np.random.default_rng(11)
    hps = rng.exponential(size=(3, 7, 301)).astype(np.float32)
    for n in (1, 10, 64, 2106):
        thr = O.quantile_thresholds(hps, n)
        for h in range(3):
            v = np.sort(hps[h].ravel().astype(np.float64))
            m = v.size
            vi = (1.0 - n / m) * (m - 1)
            i = int(np.floor(vi))
            g = vi - i
            assert thr[h] == np.float32(v[i] * (1 - g) + v[min(i + 1, m - 1)] * g)
    sub = O.quantile_thresholds(hps, 10, rows=range(2, 5))
    assert np.array_equal(sub, O.quantile_thresholds(hps[:, 2:5], 10))
