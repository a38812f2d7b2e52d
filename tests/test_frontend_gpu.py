"""GPU tests of the device front end: X = DFT(s)[:N] / Ns in FP32 (ft_conv.py:123-146).

Bar: max |X_gpu - X_fp64| <= 1e-5 of max |X_fp64| (the FOP bar) and RMS error
<= 1e-6 of the RMS spectrum, against NumPy's FP64 transform of the same float32
samples.  End to end: a time series through the front end and the engine gives
the candidate lists of the FP64 oracle on the FP64 spectrum, under the band rule
of test_gpu_parity.
"""

import numpy as np
import pytest

from fdas_amd import synth
from fdas_amd.engine import Engine, Trial
from fdas_amd.io import FrontEnd, read_tim, trim_series
from oracle import fdas_oracle as O

from test_engine_gpu import planes_of
from test_gpu_parity import assert_lists_match, bank_for
from test_io import write_tim

pytestmark = pytest.mark.gpu


def fp64_spectrum(s, n):
    return np.fft.fft(s.astype(np.float64))[:n] / (2 * n)


@pytest.mark.parametrize("log2n", [1, 3, 9, 10, 11, 12, 13, 15, 16, 19, 20, 21, 22])
def test_spectrum_matches_fp64(log2n):
    n = 1 << log2n
    w = synth.Workload("fe", n, 40 + log2n, [synth.Tone(0.31 * n, 7.0, 3, 0.05)])
    s = w.series()
    ref = fp64_spectrum(s, n)
    with FrontEnd(n) as fe:
        x = fe.run(s)
    d = np.abs(x.astype(np.complex128) - ref)
    assert d.max() <= 1e-5 * np.abs(ref).max()
    assert np.sqrt(np.mean(d ** 2)) <= 1e-6 * np.sqrt(np.mean(np.abs(ref) ** 2)) + 1e-12
    # the host oracle restatement agrees (complex64 cast of the same FP64 transform)
    if n <= 1 << 16:
        assert np.array_equal(O.spectrum(s, n), ref.astype(np.complex64))


def test_device_input_and_output():
    torch = pytest.importorskip("torch")
    n = 1 << 14
    s = synth.Workload("fe", n, 5, []).series()
    with FrontEnd(n) as fe:
        host = fe.run(s)
        out = torch.empty(n, dtype=torch.complex64, device="cuda")
        fe.run(torch.from_numpy(s).to("cuda"), out=out)
        assert np.array_equal(out.cpu().numpy(), host)
        with pytest.raises(ValueError):
            fe.run(s[:-2])


def test_time_series_file_to_candidates(tmp_path):
    """emulator-size search from a .tim file: reader -> device front end -> engine vs the FP64 oracle."""
    w = synth.emulator_workload()
    p = tmp_path / "trial.tim"
    write_tim(p, np.concatenate([w.series(), np.zeros(100, np.float32)]), tsamp=6.4e-5)
    samples, ts = read_tim(p)
    assert ts == 6.4e-5
    s = trim_series(samples, w.n_bins)
    bank = bank_for(85)
    ref = fp64_spectrum(s, w.n_bins)
    thr = synth.noise_thresholds(ref.astype(np.complex64), 8)
    with FrontEnd(w.n_bins) as fe:
        x = fe.run(s)
    with Engine(cands_per_harmonic=64, log=lambda lvl, msg: None) as eng:
        b = eng.add_bank(bank)
        (r,) = eng.run([Trial(x, b, thr, 2)])
    _, _, sel = O.run(ref.astype(np.complex64), bank, thr, 8, 64, "all")
    assert_lists_match(planes_of(r, 85), r.n_detected, sel, thr, 64, 42)
    # the injected tone (bin 5000, z = +20) is found in the fundamental plane
    _, bins = planes_of(r, 85)[0][:2]
    assert np.any(np.abs(bins - 5000) < 40)
