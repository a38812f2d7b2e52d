"""Build the native library (hipcc for gfx950) in-tree.

``python -m fdas_amd.build`` compiles ``fdas_amd/lib/libfdas_amd.so`` from
``fdas_amd/csrc``.  The library holds the HIP kernels, the host runtime behind
``include/fdas.h`` and the template-bank generator.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
LIB_DIR = PKG / "lib"
LIB = LIB_DIR / "libfdas_amd.so"

SOURCES = ["fdas_kernels.hip", "fdas_frontend.hip", "fdas_noalloc.hip", "fdas_api.cpp", "fdas_plan.cpp",
           "fdas_engine.cpp", "fdas_templates.cpp", "fdas_io.cpp"]
ARCH = os.environ.get("FDAS_OFFLOAD_ARCH", "gfx950")

# Packed FP32 instructions are kept out of every device build.
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument",
]


def hipcc() -> str:
    exe = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not Path(exe).exists():
        raise RuntimeError("hipcc not found; the FDAS library needs ROCm to build")
    return exe


def build_library(verbose: bool = False) -> Path:
    LIB_DIR.mkdir(parents=True, exist_ok=True)
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), *HIP_FLAGS, "-shared", "-o", str(tmp), *[str(CSRC / s) for s in SOURCES]]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    noise = "is not a recognized feature for this target"
    err = "\n".join(l for l in (res.stdout + res.stderr).splitlines() if noise not in l)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed ({res.returncode}):\n{err}")
    if verbose and err.strip():
        print(err, file=sys.stderr)
    os.replace(tmp, LIB)
    return LIB


FLOW_TEST_SRC = ROOT / "tests" / "cpp" / "fdas_flow_test.cpp"
FLOW_TEST = LIB_DIR / "fdas_flow_test"


def build_flow_test(verbose: bool = False) -> Path:
    """C++ driver of the reference's test flows, linked against libfdas_amd.so (host code only)."""
    cxx = shutil.which("g++") or "g++"
    cmd = [cxx, "-O2", "-std=c++17", "-Wall", f"-I{ROOT / 'include'}", str(FLOW_TEST_SRC),
           f"-L{LIB_DIR}", "-lfdas_amd", "-Wl,-rpath,$ORIGIN", "-o", str(FLOW_TEST)]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"g++ failed ({res.returncode}):\n{res.stdout}{res.stderr}")
    return FLOW_TEST


def build_all(verbose: bool = False):
    lib = build_library(verbose)
    test = build_flow_test(verbose)
    return lib, test


if __name__ == "__main__":
    print(build_all(verbose=True))
