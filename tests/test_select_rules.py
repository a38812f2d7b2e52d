"""Selection ordering rules (fdas_amd/csrc/fdas_select.h) on the host: order-preserving keys,
histogram bins and the per-thread cut search of the selection kernels, against brute force,
including planes whose detections are all negative (CPU only: built with hipcc as host code)."""

import shutil
import subprocess

import pytest

from conftest import ROOT


def test_select_cut_rules(tmp_path):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not shutil.os.path.exists(hipcc):
        pytest.skip("hipcc not available")
    exe = tmp_path / "select_cut_test"
    src = ROOT / "tests" / "cpp" / "select_cut_test.cpp"
    build = subprocess.run([hipcc, "-O2", "-std=c++17", f"-I{ROOT / 'fdas_amd' / 'csrc'}", str(src), "-o", str(exe)],
                           capture_output=True, text=True)
    assert build.returncode == 0, build.stderr[-3000:]
    run = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert run.returncode == 0, run.stdout[-3000:] + run.stderr[-2000:]
    assert "select cut test ok" in run.stdout
