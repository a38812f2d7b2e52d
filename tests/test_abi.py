"""CPU checks of the C ABI library: it loads, exports every declared entry point,
validates configurations before touching a device, and its host-only template
generator reproduces the reference bank."""

import ctypes as C
import re

import numpy as np
import pytest

import fdas_amd
from fdas_amd import _native as N
from oracle import fdas_oracle as O

HEADER = N.LIB_PATH.parents[2] / "include" / "fdas.h"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fdas_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = N.load()
    names = declared_functions()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert missing == []
    assert set(names) == set(N.SIGNATURES), "ctypes signatures out of sync with include/fdas.h"


def test_default_config_is_reference_geninfo():
    lib = N.load()
    cfg = N.FdasConfig()
    lib.fdas_default_config(C.byref(cfg))
    # gen_info.h:28-72: 85 templates, 421 taps, 2048-point FFT, 8 planes, 4096 candidates
    assert (cfg.n_templates, cfg.max_taps, cfg.tile, cfg.n_harmonics) == (85, 421, 2048, 8)
    assert cfg.cands_per_harmonic * cfg.n_harmonics == 4096


@pytest.mark.parametrize("field,value", [("n_bins", 0), ("n_bins", (1 << 22) + 1), ("n_templates", 84),
                                         ("n_templates", 257), ("tile", 3000), ("max_taps", 2048),
                                         ("n_harmonics", 0), ("n_harmonics", 17), ("cands_per_harmonic", 0),
                                         ("cands_per_harmonic", 8193)])
def test_invalid_config_rejected_before_device(field, value):
    lib = N.load()
    cfg = N.FdasConfig()
    lib.fdas_default_config(C.byref(cfg))
    cfg.n_bins = 4096
    setattr(cfg, field, value)
    msgs = []
    cb = N.LOG_FN(lambda lvl, m, u: msgs.append((lvl, m.decode())))
    ctx = C.c_void_p()
    st = lib.fdas_create(C.byref(cfg), cb, None, C.byref(ctx))
    assert st == 1  # FDAS_EINVAL
    assert not ctx
    assert msgs and msgs[-1][0] == 2 and msgs[-1][1].startswith("[ERROR]")


def test_null_arguments():
    lib = N.load()
    assert lib.fdas_create(None, N.LOG_FN(0), None, None) == 1
    assert lib.fdas_get_size(None, 0) == 0
    assert lib.fdas_last_error(None) == b""
    assert lib.fdas_status_string(7) == b"candidate spill overflow"
    assert lib.fdas_make_drift_bank(None, 1, 1, None) == 1


def test_python_mirror_reports_failure_as_false():
    f = fdas_amd.FDAS(log=lambda lvl, msg: None)
    assert f.initialise_accelerator(4096, n_templates=84) is False
    assert f.last_status == 1


def test_drift_bank_equals_reference_fixture(reference_bank):
    bank = fdas_amd.make_drift_bank(fdas_amd.reference_drifts(), 421)
    assert np.array_equal(bank, reference_bank)


@pytest.mark.parametrize("n_templates", [85, 169])
def test_drift_bank_matches_scipy_fresnel(n_templates):
    from fdas_amd import synth
    zs = synth.drift_grid(n_templates)
    ours = fdas_amd.make_drift_bank(zs, 421)
    ref = O.make_bank(zs, 421, np.complex128)
    assert np.abs(ours - ref).max() < 1e-6


def test_drift_bank_rejects_too_long_template():
    with pytest.raises(ValueError):
        fdas_amd.make_drift_bank(np.array([300.0]), 421)  # 600 taps > 421


@pytest.mark.parametrize("field,value", [("tile", 3000), ("cands_per_harmonic", 0), ("cands_per_harmonic", 8193),
                                         ("groups_in_flight", 1), ("groups_in_flight", 9), ("group_trials", 257),
                                         ("spill_per_harmonic", 100), ("rows_per_item", 1000),
                                         ("plan_schedule", 3)])
def test_invalid_engine_config_rejected_before_device(field, value):
    lib = N.load()
    cfg = N.EngineConfig()
    lib.fdas_engine_default_config(C.byref(cfg))
    assert (cfg.cands_per_harmonic, cfg.group_trials, cfg.groups_in_flight) == (4096, 8, 3)
    assert cfg.hbm_budget_bytes == 200 * 10**9
    setattr(cfg, field, value)
    msgs = []
    cb = N.LOG_FN(lambda lvl, m, u: msgs.append((lvl, m.decode())))
    eng = C.c_void_p()
    assert lib.fdas_engine_create(C.byref(cfg), cb, None, C.byref(eng)) == 1
    assert not eng
    assert msgs and msgs[-1][0] == 2 and msgs[-1][1].startswith("[ERROR]")


def test_engine_null_arguments():
    lib = N.load()
    assert lib.fdas_engine_create(None, N.LOG_FN(0), None, None) == 1
    lib.fdas_engine_destroy(None)
    lib.fdas_plan_destroy(None)
    assert lib.fdas_engine_last_error(None) == b"null engine"
    assert lib.fdas_plan_launch(None) == 1
    assert lib.fdas_engine_submit(None, None, None) == 1


def test_engine_structures_match_header():
    """ctypes layouts of the engine structures agree with the C sizes the header implies."""
    assert C.sizeof(N.EngineConfig) == 48
    assert C.sizeof(N.Request) == 8 + 6 * 4 + 16 * 4 + 8
    assert C.sizeof(N.Result) == 8 + 8 + 8 + 4 + 4 + 64 + 64 + 4 + 4 + 8
    assert C.sizeof(N.PlanInfo) == 16 + 8 + 12 + 12
    assert C.sizeof(N.ShapeInfo) == 16 + 3 * 8


def test_engine_tuning_entry_points_reject_null():
    lib = N.load()
    assert lib.fdas_engine_tune_cache(None, b"x") == 1
    assert lib.fdas_engine_shape_info(None, 1, 0, 2, None) == 1


@pytest.mark.parametrize("method", ["retrieve_candidates", "launch_resident"])
def test_candidate_arrays_must_be_contiguous(method):
    """The library writes candidate_list_sz entries from each array's first element, so a
    strided or reversed view is refused before any call reaches the device."""
    f = fdas_amd.FDAS(log=lambda lvl, msg: None)
    loc = np.zeros(64, np.uint32)[::-1]
    pwr = np.zeros(64, np.float32)
    with pytest.raises(ValueError, match="contiguous"):
        if method == "retrieve_candidates":
            f.retrieve_candidates(loc, pwr, 0)
        else:
            f.launch_resident(np.ones(8, np.float32), loc, pwr, 2, 0)
    with pytest.raises(ValueError, match="contiguous"):
        if method == "retrieve_candidates":
            f.retrieve_candidates(np.zeros(64, np.uint32), np.zeros(128, np.float32)[::2], 0)
        else:
            f.launch_resident(np.ones(8, np.float32), np.zeros(64, np.uint32), np.zeros(128, np.float32)[::2], 2, 0)


def test_engine_cancel_rejects_null_engine():
    lib = N.load()
    assert lib.fdas_engine_cancel(None, 1, None) == 1
    assert lib.fdas_status_string(8) == b"request canceled"


def test_noalloc_summing_rejects_oversize_before_device():
    """fdas_hsum_noalloc checks every size against its fixed scratch before touching a device."""
    lib = N.load()
    msgs = []
    cb = N.LOG_FN(lambda lvl, m, u: msgs.append(m.decode()))
    thr = np.ones(16, np.float32)
    dummy = C.c_void_p(0x1000)
    # 16 planes x 2^20 spill entries x 8 bytes = 128 MiB > the 32 MiB scratch
    st = lib.fdas_hsum_noalloc(dummy, 1 << 20, 85, 2, 16, thr.ctypes.data_as(N._F), 4096, 1 << 20, dummy, dummy,
                               dummy, None, cb, None)
    assert st == 1 and msgs and "scratch" in msgs[-1] and "maximum" in msgs[-1]
    st = lib.fdas_hsum_noalloc(None, 100, 85, 2, 8, thr.ctypes.data_as(N._F), 64, 0, dummy, dummy, dummy, None, cb,
                               None)
    assert st == 1
    st = lib.fdas_hsum_noalloc(dummy, 100, 84, 2, 8, thr.ctypes.data_as(N._F), 64, 0, dummy, dummy, dummy, None, cb,
                               None)
    assert st == 1 and "odd" in msgs[-1]
