"""ctypes face of the C restatement (oracle/fdas_cpu.c); test and baseline infrastructure only."""

from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
BUILD = HERE / "_build"


def build() -> None:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


def isa() -> str:
    """x86-64-v3 (AVX2 + FMA) where the host CPU has both, else x86-64-v2."""
    try:
        flags = Path("/proc/cpuinfo").read_text().split("flags", 2)[1].split("\n", 1)[0].split()
    except (OSError, IndexError):
        return "v2"
    return "v3" if {"avx2", "fma"} <= set(flags) else "v2"


def load(precision: str = "f32") -> C.CDLL:
    path = BUILD / f"libfdas_cpu_{precision}_{isa()}.so"
    if not path.exists():
        build()
    lib = C.CDLL(str(path))
    f = C.POINTER(C.c_float)
    u = C.POINTER(C.c_uint32)
    lib.fdas_cpu_fop.argtypes = [f, C.c_int, f, C.c_int, C.c_int, C.c_int, f, C.c_int]
    lib.fdas_cpu_fop.restype = C.c_int
    lib.fdas_cpu_hsum.argtypes = [f, C.c_int, C.c_int, C.c_int, f, C.c_int, u, u, f, u, C.c_int]
    lib.fdas_cpu_hsum.restype = C.c_int
    lib.fdas_cpu_threads.restype = C.c_int
    return lib


def _f(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _u(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint32))


def fop(x, bank, tile=2048, threads=0, precision="f32"):
    lib = load(precision)
    x = np.ascontiguousarray(x, dtype=np.complex64)
    bank = np.ascontiguousarray(bank, dtype=np.complex64)
    t, taps = bank.shape
    out = np.zeros((t, x.size), dtype=np.float32)
    st = lib.fdas_cpu_fop(_f(x.view(np.float32)), x.size, _f(bank.view(np.float32)), t, taps, tile, _f(out), threads)
    if st != 0:
        raise RuntimeError(f"fdas_cpu_fop failed ({st})")
    return out


def hsum(fop_plane, thresholds, n_harmonics, cap, threads=0, precision="f32"):
    """Returns (n_detected[H], list of (rows, bins, powers) per plane, sorted by key)."""
    lib = load(precision)
    fop_plane = np.ascontiguousarray(fop_plane, dtype=np.float32)
    t, n = fop_plane.shape
    thr = np.ascontiguousarray(thresholds, dtype=np.float32)
    det = np.zeros(n_harmonics, np.uint32)
    keys = np.zeros(n_harmonics * cap, np.uint32)
    pwr = np.zeros(n_harmonics * cap, np.float32)
    nout = np.zeros(n_harmonics, np.uint32)
    st = lib.fdas_cpu_hsum(_f(fop_plane), t, n, n_harmonics, _f(thr), cap, _u(det), _u(keys), _f(pwr), _u(nout),
                           threads)
    if st != 0:
        raise RuntimeError(f"fdas_cpu_hsum failed ({st})")
    planes = []
    for k in range(n_harmonics):
        m = int(nout[k])
        kk = keys[k * cap:k * cap + m]
        planes.append(((kk >> 22).astype(np.int64), (kk & 0x3FFFFF).astype(np.int64), pwr[k * cap:k * cap + m]))
    return det, planes


def threads() -> int:
    return int(load().fdas_cpu_threads())
