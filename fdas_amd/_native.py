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
                5: "FDAS_EBUSY", 6: "FDAS_ESTATE", 7: "FDAS_EOVERFLOW", 8: "FDAS_ECANCELED"}

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


MAX_HARMONICS = 16
REQ_DEVICE_SPECTRUM = 0x1
FE_DEVICE_INPUT = 0x1
FE_DEVICE_OUTPUT = 0x2
PLAN_GRAPH = 0x1


class EngineConfig(C.Structure):
    _fields_ = [
        ("tile", C.c_uint32),
        ("rows_per_item", C.c_uint32),
        ("hsum_rows", C.c_uint32),
        ("cands_per_harmonic", C.c_uint32),
        ("spill_per_harmonic", C.c_uint32),
        ("group_trials", C.c_uint32),
        ("groups_in_flight", C.c_uint32),
        ("device", C.c_int32),
        ("hbm_budget_bytes", C.c_uint64),
        ("ftc_workgroups_per_cu", C.c_uint32),
        ("plan_schedule", C.c_uint32),
    ]


class Request(C.Structure):
    _fields_ = [
        ("spectrum", C.c_void_p),
        ("n_bins", C.c_uint32),
        ("bank", C.c_uint32),
        ("n_harmonics", C.c_uint32),
        ("part", C.c_uint32),
        ("flags", C.c_uint32),
        ("pad_", C.c_uint32),
        ("thresholds", C.c_float * MAX_HARMONICS),
        ("tag", C.c_uint64),
    ]


class Result(C.Structure):
    _fields_ = [
        ("key", C.POINTER(C.c_uint32)),
        ("power", C.POINTER(C.c_float)),
        ("tag", C.c_uint64),
        ("status", C.c_int32),
        ("n_harmonics", C.c_uint32),
        ("n_detected", C.c_uint32 * MAX_HARMONICS),
        ("n_kept", C.c_uint32 * MAX_HARMONICS),
        ("tile", C.c_uint32),
        ("pad_", C.c_uint32),
        ("latency_ms", C.c_double),
    ]


class ShapeInfo(C.Structure):
    _fields_ = [
        ("tile", C.c_uint32),
        ("rows_per_item", C.c_uint32),
        ("hsum_rows", C.c_uint32),
        ("n_row_groups", C.c_uint32),
        ("forward_transforms", C.c_uint64),
        ("inverse_transforms", C.c_uint64),
        ("stored_outputs", C.c_uint64),
    ]


class EngineStats(C.Structure):
    _fields_ = [
        ("submitted", C.c_uint64),
        ("completed", C.c_uint64),
        ("groups_launched", C.c_uint64),
        ("arena_reserved_bytes", C.c_uint64),
        ("arena_in_use_bytes", C.c_uint64),
        ("arena_peak_bytes", C.c_uint64),
        ("queued", C.c_uint32),
        ("in_flight", C.c_uint32),
        ("max_group", C.c_uint32),
        ("pad_", C.c_uint32),
    ]


class PlanInfo(C.Structure):
    _fields_ = [
        ("n_trials", C.c_uint32),
        ("n_groups", C.c_uint32),
        ("n_launches", C.c_uint32),
        ("graph", C.c_uint32),
        ("region_bytes", C.c_uint64),
        ("last_ms", C.c_float),
        ("ftc_ms", C.c_float),
        ("hsum_ms", C.c_float),
        ("lagged", C.c_uint32),
        ("n_fused", C.c_uint32),
        ("fused_ms", C.c_float),
    ]


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
    "fdas_quantile_thresholds": (_I, [_P, _u32, _I, _I, _F]),
    "fdas_get_size": (C.c_uint64, [_P, _I]),
    "fdas_device_count": (_I, []),
    "fdas_device_info": (_I, [_I, C.c_char_p, C.c_size_t, C.c_char_p, C.c_size_t]),
    "fdas_platform_version": (C.c_char_p, []),
    "fdas_last_error": (C.c_char_p, [_P]),
    "fdas_status_string": (C.c_char_p, [_I]),
    "fdas_make_drift_bank": (_I, [_D, _u32, _u32, _F]),
    "fdas_reference_drifts": (_I, [_u32, C.c_double, C.c_double, C.c_double, C.c_double, _D]),
    "fdas_read_tim": (_I, [C.c_char_p, _I, _F, C.c_uint64, C.POINTER(C.c_uint64), _D]),
    "fdas_io_last_error": (C.c_char_p, []),
    "fdas_frontend_create": (_I, [_u32, _I, C.POINTER(_P)]),
    "fdas_frontend_destroy": (None, [_P]),
    "fdas_frontend_run": (_I, [_P, _P, _u32, _P]),
    "fdas_frontend_last_error": (C.c_char_p, [_P]),
    "fdas_engine_default_config": (None, [C.POINTER(EngineConfig)]),
    "fdas_engine_create": (_I, [C.POINTER(EngineConfig), LOG_FN, _P, C.POINTER(_P)]),
    "fdas_engine_destroy": (None, [_P]),
    "fdas_engine_last_error": (C.c_char_p, [_P]),
    "fdas_engine_add_bank": (_I, [_P, _F, _u32, _u32, _U]),
    "fdas_engine_prepare": (_I, [_P, _u32, _u32, _u32, _U]),
    "fdas_engine_tune_cache": (_I, [_P, C.c_char_p]),
    "fdas_engine_shape_info": (_I, [_P, _u32, _u32, _u32, C.POINTER(ShapeInfo)]),
    "fdas_plan_create": (_I, [_P, C.POINTER(Request), _u32, _u32, C.POINTER(_P)]),
    "fdas_plan_set_request": (_I, [_P, _u32, _P, _F]),
    "fdas_plan_launch": (_I, [_P]),
    "fdas_plan_wait": (_I, [_P, C.POINTER(Result)]),
    "fdas_plan_get_info": (_I, [_P, C.POINTER(PlanInfo)]),
    "fdas_plan_destroy": (None, [_P]),
    "fdas_engine_submit": (_I, [_P, C.POINTER(Request), C.POINTER(Result)]),
    "fdas_engine_poll": (_I, [_P, C.POINTER(C.POINTER(Result)), _u32, _U, C.c_int32]),
    "fdas_engine_get_stats": (_I, [_P, C.POINTER(EngineStats)]),
    "fdas_engine_cancel": (_I, [_P, C.c_uint64, _U]),
    "fdas_hsum_noalloc": (_I, [_P, _u32, _u32, _I, _u32, _F, _u32, _u32, _P, _P, _P, _P, LOG_FN, _P]),
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
    # PyTorch ships its own HIP runtime (same soname as ROCm's).  Loading it first
    # makes this library bind to that one runtime; loading ROCm's first would put
    # two HIP runtimes in the process, and the second one finds no device.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
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
