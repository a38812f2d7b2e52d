"""Pins the C restatement (oracle/fdas_cpu.c, the CPU baseline) against the NumPy oracle."""

import numpy as np
import pytest

import fdas_amd
from fdas_amd import synth
from oracle import cpu
from oracle import fdas_oracle as O


@pytest.mark.parametrize("n,tile", [(1000, 2048), (5000, 2048), (20011, 1024), (9000, 4096)])
def test_c_fop_matches_fp64_oracle(n, tile):
    x = synth.Workload("t", n, n, [synth.Tone(n * 0.3, 20.0, 4, 0.05)]).spectrum()
    bank = fdas_amd.make_drift_bank(synth.drift_grid(85), 421)
    ref = O.fop_fft(x, bank)
    assert np.abs(cpu.fop(x, bank, tile=tile) - ref).max() <= 1e-5 * ref.max()
    assert np.abs(cpu.fop(x, bank, tile=tile, precision="f64") - ref).max() <= 1e-7 * ref.max()


@pytest.mark.parametrize("t_count,n_harm", [(85, 8), (169, 16), (9, 8)])
def test_c_hsum_bit_exact_with_oracle_fp32(t_count, n_harm):
    rng = np.random.default_rng(t_count)
    fop = rng.exponential(size=(t_count, 3001)).astype(np.float32)
    thr = np.array([k + 6 * np.sqrt(k) for k in range(1, n_harm + 1)], dtype=np.float32)
    det, planes = cpu.hsum(fop, thr, n_harm, 128)
    sel = O.select(O.detect(O.harmonic_planes(fop, n_harm, dtype=np.float32), thr), 128)
    for (r, b, p), s in zip(planes, sel):
        assert np.array_equal(r, s.rows) and np.array_equal(b, s.bins)
        assert np.array_equal(p, s.powers.astype(np.float32))
    assert list(det) == [s.n_detected for s in sel]
