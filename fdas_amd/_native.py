"""ctypes binding of ``libfdas_amd.so`` (the C ABI declared in ``include/fdas.h``).

The library is built in-tree by ``python -m fdas_amd.build``.  Loading fails
loudly when it is missing: there is no CPU fallback for the device path.
"""

from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "lib" / "libfdas_amd.so"

FDAS_OK = 0
STATUS_NAMES = {0: "FDAS_OK", 1: "FDAS_EINVAL", 2: "FDAS_ENOMEM", 3: "FDAS_EDEVICE", 4: "FDAS_EUNSUPPORTED",
                5: "FDAS_EBUSY", 6: "FDAS_ESTATE", 7: "FDAS_EOVERFLOW"}

FLAG_REFERENCE_FFT_ORDER = 0x1
FLAG_SYNC_PIPELINE = 0x2

SIZE_INPUT, SIZE_TILES, SIZE_TEMPLATES, SIZE_THRESHOLDS, SIZE_FOP, SIZE_CANDIDATES = range(6)


class FdasConfig(C.Structure):
    _fields_ = [
        ("n_bins", C.c_uint32),
        ("n_templates", C.c_uint32),
        ("max_taps", C.c_uint32),
        ("tile", C.c_uint32),
        ("n_harmonics", C.c_uint32),
        ("cands_per_harmonic", C.c_uint32),
        ("spill_per_harmonic", C.c_uint32),
        ("device", C.c_int32),
        ("flags", C.c_uint32),
    ]


class FdasCandidate(C.Structure):
    _fields_ = [("harmonic", C.c_uint32), ("template_num", C.c_int32), ("bin", C.c_uint32), ("power", C.c_float)]


LOG_FN = C.CFUNCTYPE(None, C.c_int, C.c_char_p, C.c_void_p)

_P = C.c_void_p
_F = C.POINTER(C.c_float)
_U = C.POINTER(C.c_uint32)
_D = C.POINTER(C.c_double)
_I = C.c_int
_u32 = C.c_uint32

# name: (restype, argtypes)
SIGNATURES = {
    "fdas_default_config": (None, [C.POINTER(FdasConfig)]),
    "fdas_create": (_I, [C.POINTER(FdasConfig), LOG_FN, _P, C.POINTER(_P)]),
    "fdas_destroy": (None, [_P]),
    "fdas_upload_templates": (_I, [_P, _F, _I]),
    "fdas_upload_template_taps": (_I, [_P, _F, _u32, _I]),
    "fdas_upload_thresholds": (_I, [_P, _F, _I]),
    "fdas_perform_input_tiling": (_I, [_P, _F, _I]),
    "fdas_set_input_device": (_I, [_P, _P, _I]),
    "fdas_perform_ft_convolution": (_I, [_P, _I, _I]),
    "fdas_perform_harmonic_summing": (_I, [_P, _F, _I, _I]),
    "fdas_launch": (_I, [_P, _F, _F, _U, _F, _I, _I]),
    "fdas_launch_resident": (_I, [_P, _F, _U, _F, _I, _I]),
    "fdas_wait": (_I, [_P, _I]),
    "fdas_end_pipeline": (_I, [_P, _I]),
    "fdas_retrieve_tiles": (_I, [_P, _F, _I]),
    "fdas_retrieve_fop": (_I, [_P, _F, _I]),
    "fdas_inject_fop": (_I, [_P, _F, _I]),
    "fdas_retrieve_candidates": (_I, [_P, _U, _F, _I]),
    "fdas_retrieve_candidates_wide": (_I, [_P, C.POINTER(FdasCandidate), _u32, _U, _U, _I]),
    "fdas_retrieve_harmonic_plane": (_I, [_P, _u32, _F, _I, _I]),
    "fdas_print_stats": (None, [_P, _I, _I]),
    "fdas_print_events": (None, [_P, _I]),
    "fdas_get_stage_times": (_I, [_P, _I, _F]),
    "fdas_get_size": (C.c_uint64, [_P, _I]),
    "fdas_device_count": (_I, []),
    "fdas_device_info": (_I, [_I, C.c_char_p, C.c_size_t, C.c_char_p, C.c_size_t]),
    "fdas_platform_version": (C.c_char_p, []),
    "fdas_last_error": (C.c_char_p, [_P]),
    "fdas_status_string": (C.c_char_p, [_I]),
    "fdas_make_drift_bank": (_I, [_D, _u32, _u32, _F]),
    "fdas_reference_drifts": (_I, [_u32, C.c_double, C.c_double, C.c_double, C.c_double, _D]),
}

_lib = None


def load() -> C.CDLL:
    """Load the native library once; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m fdas_amd.build` "
                           "(the FDAS device path has no CPU fallback)")
    lib = C.CDLL(str(LIB_PATH))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def fptr(a):
    return a.ctypes.data_as(_F)


def uptr(a):
    return a.ctypes.data_as(_U)


def dptr(a):
    return a.ctypes.data_as(_D)
