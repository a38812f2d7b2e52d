"""GPU parity of the HIP path (through the C ABI) against the FP64 oracle.

Bars (from the reference's tests, tightened):
  * tiles: |delta| < 1e-5 absolute per component (main_gtest.cpp:56,173-181);
  * FOP and harmonic planes: max |delta| <= 1e-5 of the plane's maximum, FP64 oracle;
  * candidates: sorted (template, bin, harmonic) lists identical except entries
    within 1e-4 relative of the plane's cut (its threshold, or the power of the
    last kept entry when the cap binds); powers within 1e-5 of the largest;
  * harmonic summing fed the oracle's own float32 FOP: bit-exact planes and lists
    (same FP32 additions in the same order as hsum.py:93-120).
"""

import numpy as np
import pytest

import fdas_amd
from fdas_amd import synth
from fdas_amd.host import FDAS
from oracle import fdas_oracle as O

pytestmark = pytest.mark.gpu

PARTS = {FDAS.NegativeAccelerations: "neg", FDAS.PositiveAccelerations: "pos", FDAS.AllAccelerations: "all"}
FOP_TOL = 1e-5
BAND = 1e-4


def make(n_bins, **kw):
    f = FDAS(log=lambda lvl, msg: None)
    assert f.initialise_accelerator(n_bins, **kw), f.last_error
    return f


def bank_for(n_templates=85, max_taps=421):
    return fdas_amd.make_drift_bank(synth.drift_grid(n_templates), max_taps)


def spectrum(n, seed, f0_frac=0.3, z=20.0, amp=0.05):
    return synth.Workload("t", n, seed, [synth.Tone(max(1.0, n * f0_frac), z, 4, amp)]).spectrum()


def rel_err(gpu, ref):
    scale = float(np.max(ref))
    return float(np.max(np.abs(gpu.astype(np.float64) - ref))) / (scale if scale > 0 else 1.0)


def wide_by_plane(f, ab, n_harmonics):
    h, t, b, p, det = f.retrieve_candidates_wide(ab)
    out = []
    for k in range(1, n_harmonics + 1):
        m = h == k
        out.append((t[m], b[m], p[m]))
    return out, det


def assert_lists_match(gpu_planes, gpu_det, oracle_sel, thresholds, cap, c, exact=False):
    for k, (sel, (gt, gb, gp)) in enumerate(zip(oracle_sel, gpu_planes), start=1):
        # each plane's list is sorted by (template, bin)
        assert list(zip(gt.tolist(), gb.tolist())) == sorted(zip(gt.tolist(), gb.tolist()))
        o = {(int(r) - c, int(b)): float(p) for r, b, p in zip(sel.rows, sel.bins, sel.powers)}
        g = {(int(t), int(b)): float(p) for t, b, p in zip(gt, gb, gp)}
        if exact:
            assert gpu_det[k - 1] == sel.n_detected
            assert g == {key: float(np.float32(v)) for key, v in o.items()}
            continue
        cut = float(thresholds[k - 1])
        if sel.n_detected > cap:
            cut = max(cut, min(o.values()))
        for key in set(o) ^ set(g):
            p = o.get(key, g.get(key))
            assert abs(p - cut) <= BAND * cut, f"plane {k}: {key} power {p} not within {BAND} of cut {cut}"
        if o:
            top = max(o.values())
            for key in set(o) & set(g):
                assert abs(g[key] - o[key]) <= FOP_TOL * top
        assert abs(int(gpu_det[k - 1]) - sel.n_detected) <= max(2, int(0.001 * sel.n_detected))


# -- tiles -----------------------------------------------------------------

@pytest.mark.parametrize("tile", [1024, 2048, 4096])
@pytest.mark.parametrize("p4", [False, True])
def test_tiles(tile, p4):
    n = 9000
    x = spectrum(n, 21)
    f = make(n, tile=tile, reference_fft_order=p4)
    assert f.perform_input_tiling(x, FDAS.A), f.last_error
    tiles = np.zeros(f.get_tiles_sz(), np.complex64)
    assert f.retrieve_tiles(tiles, FDAS.A), f.last_error
    ref = O.tiles_transformed(x, O.geometry(n, tile))
    if p4:
        ref = O.to_fft_order(ref)
    d = tiles.reshape(ref.shape) - ref
    assert np.abs(d.real).max() < 1e-5 and np.abs(d.imag).max() < 1e-5
    f.close()


# -- FOP --------------------------------------------------------------------

@pytest.mark.parametrize("n", [1, 419, 1628, 1629, 5000, 20011])
@pytest.mark.parametrize("tile", [1024, 2048, 4096])
def test_fop_all_rows(n, tile):
    x = spectrum(n, n + tile)
    bank = bank_for()
    f = make(n, tile=tile)
    assert f.upload_template_taps(bank, FDAS.B), f.last_error
    assert f.perform_input_tiling(x, FDAS.B), f.last_error
    assert f.perform_ft_convolution(FDAS.AllAccelerations, FDAS.B), f.last_error
    fop = np.zeros(f.get_fop_sz(), np.float32)
    assert f.retrieve_FOP(fop, FDAS.B), f.last_error
    ref = O.fop_fft(x, bank)
    assert rel_err(fop.reshape(85, n), ref) <= FOP_TOL
    f.close()


@pytest.mark.parametrize("part", [FDAS.NegativeAccelerations, FDAS.PositiveAccelerations])
def test_fop_parts_touch_only_their_rows(part):
    n = 7000
    x = spectrum(n, 3)
    bank = bank_for()
    f = make(n)
    assert f.upload_template_taps(bank, FDAS.A)
    assert f.perform_input_tiling(x, FDAS.A)
    assert f.perform_ft_convolution(part, FDAS.A), f.last_error
    fop = np.zeros(f.get_fop_sz(), np.float32)
    assert f.retrieve_FOP(fop, FDAS.A)
    fop = fop.reshape(85, n)
    rows = list(O.part_rows(85, PARTS[part]))
    ref = O.fop_fft(x, bank, rows)
    assert rel_err(fop[rows], ref[rows]) <= FOP_TOL
    other = [r for r in range(85) if r not in rows]
    assert np.all(fop[other] == 0)
    f.close()


def test_reference_bank_in_fft_order(reference_bank, reference_bank_ft_p4):
    """The reference's own template file (85 x 2048, FFT order p=4) through upload_templates."""
    n = 12000
    x = spectrum(n, 4)
    f = make(n, reference_fft_order=True)
    assert f.upload_templates(reference_bank_ft_p4, FDAS.A), f.last_error
    assert f.perform_input_tiling(x, FDAS.A)
    assert f.perform_ft_convolution(FDAS.AllAccelerations, FDAS.A)
    fop = np.zeros(f.get_fop_sz(), np.float32)
    assert f.retrieve_FOP(fop, FDAS.A)
    ref = O.fop_fft(x, reference_bank)
    # the file holds complex64 spectra: allow their rounding (1.2e-7 abs per bin)
    assert rel_err(fop.reshape(85, n), ref) <= FOP_TOL
    assert np.allclose(fop.reshape(85, n)[42], np.abs(x.astype(np.complex128)) ** 2, rtol=1e-4,
                       atol=1e-5 * ref.max())
    f.close()


# -- golden fixture -----------------------------------------------------------

def test_golden_small(golden_small):
    g = golden_small
    n = g["spectrum"].size
    f = make(n, cands_per_harmonic=64)
    assert f.upload_template_taps(g["bank"], FDAS.A)
    assert f.perform_input_tiling(g["spectrum"], FDAS.A)
    assert f.perform_ft_convolution(FDAS.AllAccelerations, FDAS.A)
    fop = np.zeros(f.get_fop_sz(), np.float32)
    assert f.retrieve_FOP(fop, FDAS.A)
    assert rel_err(fop.reshape(85, n), g["fop"].astype(np.float64)) <= FOP_TOL
    assert f.perform_harmonic_summing(g["thresholds"], FDAS.AllAccelerations, FDAS.A), f.last_error
    planes, det = wide_by_plane(f, FDAS.A, 8)
    sel = [O.PlaneCandidates(g[f"rows_{k}"], g[f"bins_{k}"], g[f"powers_{k}"], int(g["n_detected"][k - 1]))
           for k in range(1, 9)]
    assert_lists_match(planes, det, sel, g["thresholds"], 64, 42)
    f.close()


# -- harmonic summing fed a known FOP ---------------------------------------------

@pytest.mark.parametrize("part", list(PARTS))
@pytest.mark.parametrize("n_templates,n_harmonics", [(85, 8), (169, 16), (9, 8)])
@pytest.mark.parametrize("n", [6007, 6008])  # per-bin loads / 16-byte loads (n % 8 == 0)
def test_hsum_bit_exact_on_injected_fop(part, n_templates, n_harmonics, n):
    rng = np.random.default_rng(n_templates + part)
    fop = rng.exponential(1.0, size=(n_templates, n)).astype(np.float32)
    fop[n_templates // 2 + 3, 1234] = 60.0  # a strong line
    f = make(n, n_templates=n_templates, n_harmonics=n_harmonics, cands_per_harmonic=256)
    assert f.inject_FOP(fop, FDAS.A), f.last_error
    thr = np.array([k + 6.0 * np.sqrt(k) for k in range(1, n_harmonics + 1)], dtype=np.float32)
    ok = f.perform_harmonic_summing(thr, part, FDAS.A)
    assert ok, f.last_error
    rows = O.part_rows(n_templates, PARTS[part])
    hps = O.harmonic_planes(fop, n_harmonics, rows, dtype=np.float32)
    for k in (1, 2, n_harmonics):
        plane = f.retrieve_harmonic_plane(k, part, FDAS.A)
        assert np.array_equal(plane[list(rows)], hps[k - 1][list(rows)])
    sel = O.select(O.detect(hps, thr, rows), 256)
    planes, det = wide_by_plane(f, FDAS.A, n_harmonics)
    assert_lists_match(planes, det, sel, thr, 256, n_templates // 2, exact=True)
    f.close()


# -- reference candidate layout and validator semantics ------------------------

def reference_validate(loc, pwr, loc_ref, pwr_ref, tol=1e-5):
    """main_gtest.cpp:192-228: drop invalid slots; 1 <= |list| <= |ref|; every entry in ref, |dp| < tol."""
    keep = loc != O.invalid_location()
    loc, pwr = loc[keep], pwr[keep]
    assert 1 <= loc.size <= loc_ref.size
    where = {int(v): i for i, v in enumerate(loc_ref)}
    for v, p in zip(loc, pwr):
        assert int(v) in where
        assert abs(p - pwr_ref[where[int(v)]]) < tol * max(1.0, abs(pwr_ref[where[int(v)]]))


def test_reference_candidate_layout():
    n = 6000
    rng = np.random.default_rng(11)
    fop = rng.exponential(1.0, size=(85, n)).astype(np.float32)
    f = make(n)  # 512 per plane -> 4096-entry lists
    assert f.get_candidate_list_sz() == 4096
    assert f.inject_FOP(fop, FDAS.B)
    thr = np.array([k + 6.0 * np.sqrt(k) for k in range(1, 9)], dtype=np.float32)
    assert f.perform_harmonic_summing(thr, FDAS.NegativeAccelerations, FDAS.B), f.last_error
    loc = np.zeros(4096, np.uint32)
    pwr = np.zeros(4096, np.float32)
    assert f.retrieve_candidates(loc, pwr, FDAS.B), f.last_error
    rows = O.part_rows(85, "neg")
    hps = O.harmonic_planes(fop, 8, rows, dtype=np.float32)
    det = O.detect(hps, thr, rows)
    sel = O.select(det, 512)
    ref_loc, ref_pwr = O.reference_lists(sel, 512)
    assert np.array_equal(loc, ref_loc) and np.array_equal(pwr, ref_pwr)
    # the full (uncapped) reference list, as hsum.py writes it
    full = O.select(det, 10 ** 9)
    all_loc = np.concatenate([O.encode_location(k + 1, s.rows - 42, s.bins) for k, s in enumerate(full)])
    all_pwr = np.concatenate([s.powers for s in full]).astype(np.float32)
    reference_validate(loc, pwr, all_loc, all_pwr)
    f.close()


def test_wide_only_configuration_rejects_reference_layout():
    n = 3000
    f = make(n, n_templates=169, n_harmonics=16, cands_per_harmonic=32)
    fop = np.random.default_rng(1).exponential(size=(169, n)).astype(np.float32)
    assert f.inject_FOP(fop, FDAS.A)
    thr = np.full(16, 30.0, dtype=np.float32)
    assert f.perform_harmonic_summing(thr, FDAS.AllAccelerations, FDAS.A)
    loc = np.zeros(f.get_candidate_list_sz(), np.uint32)
    pwr = np.zeros_like(loc, dtype=np.float32)
    assert f.retrieve_candidates(loc, pwr, FDAS.A) is False
    assert f.last_status == 4  # FDAS_EUNSUPPORTED
    f.close()


@pytest.mark.parametrize("n", [4096, 4001])
def test_hsum_dense_detections_bit_exact(n):
    """Thresholds at half the plane mean: most bins fire, so a wave row holds more
    hits than its LDS stage and they go straight to the spill list (16-byte-load
    form at n = 4096, per-bin loads at n = 4001)."""
    rng = np.random.default_rng(n)
    fop = rng.exponential(1.0, size=(85, n)).astype(np.float32)
    f = make(n, cands_per_harmonic=256, spill_per_harmonic=85 * n)  # room for every bin
    assert f.inject_FOP(fop, FDAS.A), f.last_error
    thr = np.array([0.5 * k for k in range(1, 9)], dtype=np.float32)
    assert f.perform_harmonic_summing(thr, FDAS.AllAccelerations, FDAS.A), f.last_error
    rows = O.part_rows(85, "all")
    hps = O.harmonic_planes(fop, 8, rows, dtype=np.float32)
    det = O.detect(hps, thr, rows)
    assert min(d.powers.size for d in det) > 85 * n // 4  # dense in every plane
    sel = O.select(det, 256)
    planes, gdet = wide_by_plane(f, FDAS.A, 8)
    assert_lists_match(planes, gdet, sel, thr, 256, 42, exact=True)
    f.close()


# -- drivers of the reference test (main_gtest.cpp:315-437) ---------------------

def _blocking_lists(x, bank, thr, part):
    f = make(x.size)
    assert f.upload_template_taps(bank, FDAS.A)
    assert f.perform_input_tiling(x, FDAS.A)
    assert f.perform_ft_convolution(part, FDAS.A)
    assert f.perform_harmonic_summing(thr, part, FDAS.A)
    loc = np.zeros(f.get_candidate_list_sz(), np.uint32)
    pwr = np.zeros_like(loc, dtype=np.float32)
    assert f.retrieve_candidates(loc, pwr, FDAS.A)
    f.close()
    return loc, pwr


@pytest.mark.parametrize("sync", [False, True])
def test_pipelined_launches_match_blocking(sync):
    n = 16000
    x = spectrum(n, 8)
    bank = bank_for()
    thr = synth.noise_thresholds(x, 8)
    expect = {p: _blocking_lists(x, bank, thr, p) for p in (FDAS.PositiveAccelerations, FDAS.NegativeAccelerations)}
    f = make(n, sync_pipeline=sync)
    for ab in (FDAS.A, FDAS.B):
        assert f.upload_template_taps(bank, ab)
    which = {FDAS.A: FDAS.PositiveAccelerations, FDAS.B: FDAS.NegativeAccelerations}
    locs = {ab: np.zeros(f.get_candidate_list_sz(), np.uint32) for ab in (FDAS.A, FDAS.B)}
    pwrs = {ab: np.zeros(f.get_candidate_list_sz(), np.float32) for ab in (FDAS.A, FDAS.B)}
    n_runs = 6
    for i in range(n_runs + 1):
        ab, prev = (FDAS.A, FDAS.B) if i % 2 == 0 else (FDAS.B, FDAS.A)
        if i < n_runs:
            assert f.launch(x, thr, locs[ab], pwrs[ab], which[ab], ab), f.last_error
        else:
            assert f.end_pipeline(ab)
        if i > 0:
            assert f.wait(prev), f.last_error
            loc, pwr = expect[which[prev]]
            assert np.array_equal(locs[prev], loc) and np.array_equal(pwrs[prev], pwr)
            assert f.stage_times(prev)[5] > 0 and f.stage_times(prev)[1] > 0
    f.close()


def test_serial_launches_alternate_parts():
    n = 9000
    x = spectrum(n, 9)
    bank = bank_for()
    thr = synth.noise_thresholds(x, 8)
    f = make(n)
    assert f.upload_template_taps(bank, FDAS.A)
    loc = np.zeros(f.get_candidate_list_sz(), np.uint32)
    pwr = np.zeros(f.get_candidate_list_sz(), np.float32)
    for i in range(4):
        part = FDAS.PositiveAccelerations if i % 2 == 0 else FDAS.NegativeAccelerations
        assert f.launch(x, thr, loc, pwr, part, FDAS.A), f.last_error
        assert f.wait(FDAS.A), f.last_error
        e_loc, e_pwr = _blocking_lists(x, bank, thr, part)
        assert np.array_equal(loc, e_loc) and np.array_equal(pwr, e_pwr)
    f.close()


# -- failure handling ---------------------------------------------------------------

def test_spill_overflow_is_reported():
    n = 4000
    f = make(n, spill_per_harmonic=100, cands_per_harmonic=16)
    fop = np.random.default_rng(2).exponential(size=(85, n)).astype(np.float32)
    assert f.inject_FOP(fop, FDAS.A)
    thr = np.zeros(8, dtype=np.float32)  # everything is a detection
    assert f.perform_harmonic_summing(thr, FDAS.AllAccelerations, FDAS.A) is False
    assert f.last_status == 7 and "spill" in f.last_error
    f.close()


def test_calls_out_of_order_fail_cleanly():
    f = make(2048)
    assert f.perform_ft_convolution(FDAS.AllAccelerations, FDAS.A) is False
    assert f.last_status == 6
    x = spectrum(2048, 1)
    assert f.perform_input_tiling(x, FDAS.A)
    assert f.perform_ft_convolution(FDAS.AllAccelerations, FDAS.A) is False  # still no templates
    assert f.wait(FDAS.B) is False
    assert f.perform_harmonic_summing(None, FDAS.AllAccelerations, FDAS.A) is False  # no thresholds
    f.close()


# -- full single-trial size ---------------------------------------------------------

@pytest.mark.slow
def test_full_size_single_trial():
    """N = 2^21, 85 templates, 8 harmonics, 4096 candidates per plane."""
    w = synth.full_workload()
    x = w.spectrum()
    bank = bank_for()
    thr = synth.noise_thresholds(x, 8)
    f = make(w.n_bins, cands_per_harmonic=4096)
    assert f.upload_template_taps(bank, FDAS.A)
    assert f.perform_input_tiling(x, FDAS.A)
    assert f.perform_ft_convolution(FDAS.AllAccelerations, FDAS.A)
    assert f.perform_harmonic_summing(thr, FDAS.AllAccelerations, FDAS.A), f.last_error
    fop = np.zeros(f.get_fop_sz(), np.float32)
    assert f.retrieve_FOP(fop, FDAS.A)
    fop = fop.reshape(85, w.n_bins)
    ref = O.fop_fft(x, bank)
    assert rel_err(fop, ref) <= FOP_TOL
    sel, kept = O.detect_streaming(ref, 8, thr, keep_planes=(8,))
    hp8 = f.retrieve_harmonic_plane(8, FDAS.AllAccelerations, FDAS.A)
    assert rel_err(hp8, kept[8]) <= FOP_TOL
    del kept, hp8
    sel = O.select(sel, 4096)
    planes, det = wide_by_plane(f, FDAS.A, 8)
    assert_lists_match(planes, det, sel, thr, 4096, 42)
    # the injected tone (fundamental at bin 300,000, drifting by z = -36) is detected in HP_1
    assert np.any(np.abs(planes[0][1] - 300000) < 100)
    f.close()


@pytest.mark.parametrize("cap", [100, 4096])
def test_selection_with_massive_ties(cap):
    """Equal powers everywhere: the boundary bin exceeds the LDS pool and the exact
    radix-select path decides; ties go to the smaller (template, bin)."""
    n = 5000
    fop = np.full((85, n), 0.25, dtype=np.float32)
    fop[:, ::7] = 0.5
    f = make(n, cands_per_harmonic=cap)
    assert f.inject_FOP(fop, FDAS.A)
    thr = np.array([0.3 * k for k in range(1, 9)], dtype=np.float32)
    assert f.perform_harmonic_summing(thr, FDAS.AllAccelerations, FDAS.A), f.last_error
    rows = O.part_rows(85, "all")
    hps = O.harmonic_planes(fop, 8, rows, dtype=np.float32)
    sel = O.select(O.detect(hps, thr, rows), cap)
    planes, det = wide_by_plane(f, FDAS.A, 8)
    assert_lists_match(planes, det, sel, thr, cap, 42, exact=True)
    f.close()


# -- the reference's mock-up thresholds (hsum.py:125-131) ------------------------

@pytest.mark.parametrize("part", list(PARTS))
@pytest.mark.parametrize("n_templates,n_harmonics,n_cands", [(85, 8, 64), (9, 4, 10), (169, 16, 500)])
def test_quantile_thresholds_on_device(part, n_templates, n_harmonics, n_cands):
    n = 6007
    rng = np.random.default_rng(7 + part + n_templates)
    fop = rng.exponential(1.0, size=(n_templates, n)).astype(np.float32)
    fop[n_templates // 2, 3000] = 200.0  # a line that dominates a few order statistics
    f = make(n, n_templates=n_templates, n_harmonics=n_harmonics, cands_per_harmonic=256)
    assert f.inject_FOP(fop, FDAS.A), f.last_error
    thr = f.quantile_thresholds(n_cands, part, FDAS.A)
    assert thr is not None, f.last_error
    rows = O.part_rows(n_templates, PARTS[part])
    hps = O.harmonic_planes(fop, n_harmonics, rows, dtype=np.float32)
    ref = O.quantile_thresholds(hps, n_cands, rows)
    assert np.array_equal(thr, ref)
    # about n candidates per plane: the threshold lies between the n-th and (n+1)-th largest value
    for k in range(n_harmonics):
        assert n_cands - 1 <= int(np.sum(hps[k][list(rows)] > thr[k])) <= n_cands
    f.close()
