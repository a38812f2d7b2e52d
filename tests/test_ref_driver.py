"""The reference's own test driver (fdas/src/host/main_gtest.cpp) against this library.

* CPU: the reference test driver compiles unchanged from /root/reference, with its vendored googletest and
  libnpy, against tests/cpp/refshim/FDAS.h (a drop-in FDAS.h over include/FDAS.hpp) and links
  with libfdas_amd.so (oracle/Makefile, target ref-driver).  Skipped where /root/reference is
  absent.
* GPU: the reference test driver built that way (oracle/_ref/fdas_gtest, built by __graft_entry__.build())
  runs its serial and pipelined flows (three buffer modes each, main_gtest.cpp:315-437) on
  test vectors written in its own layout at the emulator size (N = 32,768, the reference's
  85-template bank), and its validator (main_gtest.cpp:192-228) accepts every list.
"""

import os
import shutil
import subprocess

import numpy as np
import pytest

from conftest import ROOT
from fdas_amd import synth

from ref_vectors import driver_tree, write_vectors

REF = "/root/reference"
DRIVER = ROOT / "oracle" / "_ref" / "fdas_gtest"


@pytest.mark.skipif(not os.path.isdir(os.path.join(REF, "fdas", "src", "host")), reason="reference tree absent")
def test_reference_driver_compiles_against_the_shim(tmp_path):
    if not (ROOT / "fdas_amd" / "lib" / "libfdas_amd.so").exists():
        pytest.skip("library not built")
    res = subprocess.run(["make", "-s", "-C", str(ROOT / "oracle"), "ref-driver", f"REF_OUT={tmp_path}"],
                         capture_output=True, text=True, timeout=600)
    assert res.returncode == 0, res.stdout[-2000:] + res.stderr[-4000:]
    assert (tmp_path / "fdas_gtest").exists()


def test_reference_vectors_layout(tmp_path):
    """The vector writer's files load with the dtypes and shapes the reference test driver asserts."""
    x = synth.emulator_workload().spectrum()[:6000]
    bank = np.load(ROOT / "tests" / "golden" / "fdas_templates_85_350.npy", allow_pickle=False)
    info = write_vectors(tmp_path, x, bank)
    assert np.load(tmp_path / "input.npy").dtype == np.complex64
    assert np.load(tmp_path / "fop_ref.npy").shape == (85, 6000)
    assert np.load(tmp_path / "thresholds.npy").dtype == np.float32
    loc = np.load(tmp_path / "det_loc_ref.npy")
    assert loc.dtype == np.uint32 and loc.size == info["n_detections"] > 0
    assert np.load(tmp_path / "input_tiled_p4_ref.npy").shape[1] == 2048


@pytest.mark.gpu
def test_reference_driver_flows_pass(tmp_path):
    if not DRIVER.exists():
        pytest.skip("oracle/_ref/fdas_gtest not built (needs the reference tree at build time)")
    x = synth.emulator_workload().spectrum()
    bank = np.load(ROOT / "tests" / "golden" / "fdas_templates_85_350.npy", allow_pickle=False)
    vec = tmp_path / "vectors"
    info = write_vectors(vec, x, bank)
    assert info["n_detections"] > 8
    cwd = driver_tree(tmp_path)
    flows = "*FDAS_serial_set*:*FDAS_serial_stage*:*FDAS_pipelined_set*:*FDAS_pipelined_stage*:*FDAS_pipelined_crossed*"
    print(f"reference driver: {DRIVER} -N 3 {vec}", flush=True)
    res = subprocess.run([str(DRIVER), "-N", "3", str(vec), f"--gtest_filter={flows}"], cwd=cwd,
                         capture_output=True, text=True, timeout=240)
    out = res.stdout + res.stderr
    assert res.returncode == 0, out[-6000:]
    assert "[  PASSED  ] 5 tests" in out, out[-6000:]
    assert "INVALID" not in out
