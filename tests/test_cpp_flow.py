"""The reference's host test flows (main_gtest.cpp) run in C++ against the drop-in class.

The oracle writes a test-vector directory in the reference's layout
(input.npy, thresholds.npy, fop_ref.npy, det_loc_ref.npy, det_pwr_ref.npy,
ft_conv.py / hsum.py naming) and the C++ driver tests/cpp/fdas_flow_test.cpp
uploads the reference's own template file in FFT order.
"""

import subprocess

import numpy as np
import pytest

from fdas_amd import build as B
from fdas_amd import synth
from oracle import fdas_oracle as O

from conftest import GOLDEN


def write_vectors(d, n=12000, seed=31):
    bank = np.load(GOLDEN / "fdas_templates_85_350.npy", allow_pickle=False)
    x = synth.Workload("flow", n, seed, [synth.Tone(n * 0.27, 60.0, 4, 0.06)]).spectrum()
    fop = O.fop_fft(x, bank)
    hps = O.harmonic_planes(fop, 8)
    # hsum.py's mock thresholds (~64 candidates per plane), moved to the middle of
    # the gap around them so that no FP64 value sits within 1e-4 of a threshold
    thr = O.quantile_thresholds(hps, 64).astype(np.float64)
    for h in range(8):
        v = np.sort(hps[h].reshape(-1))
        i = np.searchsorted(v, thr[h])
        lo, hi = v[max(i - 1, 0)], v[min(i, v.size - 1)]
        thr[h] = 0.5 * (lo + hi)
        assert min(abs(lo - thr[h]), abs(hi - thr[h])) > 2e-4 * thr[h] or hi == lo
    thr = thr.astype(np.float32)
    det = O.select(O.detect(hps, thr), 10 ** 9)
    loc = np.concatenate([O.encode_location(k + 1, s.rows - 42, s.bins) for k, s in enumerate(det)])
    pwr = np.concatenate([s.powers for s in det]).astype(np.float32)
    np.save(d / "input.npy", x)
    np.save(d / "thresholds.npy", thr)
    np.save(d / "fop_ref.npy", fop.astype(np.float32))
    np.save(d / "det_loc_ref.npy", loc.astype(np.uint32))
    np.save(d / "det_pwr_ref.npy", pwr)
    return loc.size


def test_flow_binary_is_built():
    assert B.FLOW_TEST.exists(), "build the C++ flow driver with python -m fdas_amd.build"
    r = subprocess.run([str(B.FLOW_TEST)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "usage" in r.stderr


@pytest.mark.gpu
def test_reference_flows_in_cpp(tmp_path):
    n_ref = write_vectors(tmp_path)
    assert n_ref > 100
    r = subprocess.run([str(B.FLOW_TEST), "-t", str(GOLDEN / "fdas_templates_85_350_ft_p4.npy"), "-N", "4",
                        str(tmp_path)], capture_output=True, text=True, timeout=180)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr[-3000:]
    assert r.stdout.count("  OK  ") == 11 and "FAIL" not in r.stdout
