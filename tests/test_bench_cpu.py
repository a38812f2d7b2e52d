"""bench.py's arithmetic and contract on the CPU: the multi-GPU refusal, the algorithmic byte counts
behind the roofline, the VALU issue-time model and the amd-smi parser of the power form."""

import importlib.util
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def bench():
    spec = importlib.util.spec_from_file_location("bench_under_test", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


def test_algorithmic_bytes(bench):
    n, t = 1 << 21, 85
    # per trial: spectrum in (8N) + FOP out (4TN); template spectra (8TS) once per launch
    assert bench.ftc_bytes(n, t, 1024) == 8 * n + 4 * t * n + 8 * t * 1024
    # the 8-trial launch of the profiled batched shape
    assert bench.ftc_bytes(n, t, 1024, 8) == 5_839_167_488
    # summing: FOP read once per trial + 8 bytes per candidate
    assert bench.hsum_bytes(n, t, 100, 8) == 8 * 4 * t * n + 800
    # whole path per trial: spectrum in, FOP written and read back
    assert bench.path_bytes(n, t) == 8 * n + 8 * t * n


def test_valu_issue_model(bench):
    k = "ftconv1w_kernel"
    n = bench.MEASURED_VALU[(k, 8)]
    v = bench.valu_issue(k, 8, 2.0)
    # one wave64 VALU instruction per 4 cycles on each of 256 x 4 SIMDs at 2.4 GHz
    assert v["issue_ms"] == round(n * 4 / 1024 / 2.4e9 * 1e3, 4)
    assert v["frac_of_span"] == round(n * 4 / 1024 / 2.4e9 / 2.0e-3, 3)
    # at the measured clock the issue time scales by 2.4 / clock
    ghz = bench.MEASURED_CLOCK_GHZ[k]
    assert v["issue_ms_at_clock"] == round(n * 4 / 1024 / (ghz * 1e9) * 1e3, 4)
    assert v["clock_ghz"] == ghz


SMI_SAMPLE = """GPU: 0
    POWER:
        SOCKET_POWER: 1393 W
        GFX_VOLTAGE: N/A
    CLOCK:
        GFX_0:
            CLK: 2093 MHz
            MIN_CLK: 500 MHz
            MAX_CLK: 2400 MHz
        GFX_1:
            CLK: 2160 MHz
            MIN_CLK: 500 MHz
            MAX_CLK: 2400 MHz
        MEM_0:
            CLK: 2000 MHz
"""


def test_amd_smi_parser(bench):
    pw, mean, lo, hi = bench.parse_smi_metric(SMI_SAMPLE)
    assert pw == 1393.0 and (lo, hi) == (2093, 2160) and mean == pytest.approx(2126.5)
    # the memory clock is not a shader clock; a reply without power or clocks gives None
    assert bench.parse_smi_metric(SMI_SAMPLE.replace("SOCKET_POWER: 1393 W", "SOCKET_POWER: N/A")) is None
    assert bench.parse_smi_metric("GPU: 0\n    POWER:\n        SOCKET_POWER: 250 W\n") is None
