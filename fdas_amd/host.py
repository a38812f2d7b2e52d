"""Python mirror of the reference host class ``FDAS`` over the C ABI.

Method names, argument meaning and ``bool`` returns follow
``fdas/src/host/FDAS.h:36-92`` so that a test written against the reference
flow (``fdas/src/host/main_gtest.cpp``) reads the same here.  The error text of
a failed call is in ``FDAS.last_error`` and was written to the log, as the
reference writes ``[ERROR] ...`` lines to its ``std::ostream``
(``FDAS.cpp:34-52``).
"""

from __future__ import annotations

import ctypes as C
import sys

import numpy as np

from . import _native as N


def _c64(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.complex64)
    if shape is not None and a.size != int(np.prod(shape)):
        raise ValueError(f"expected {int(np.prod(shape))} complex values, got {a.size}")
    return a


class FDAS:
    # FDAS::FOPPart, FDAS::BufferSet, FDAS::BufferMode (FDAS.h:38-40)
    NegativeAccelerations, PositiveAccelerations, AllAccelerations = 0, 1, 2
    A, B = 0, 1
    PerSet, PerStage, Crossed = 0, 1, 2

    def __init__(self, log=None):
        """``log`` is a text stream (default ``sys.stderr``) or a callable(level, message)."""
        self._lib = N.load()
        self._ctx = C.c_void_p()
        self._log_target = sys.stderr if log is None else log
        self._log_cb = N.LOG_FN(self._on_log)
        self._keep = {0: (), 1: ()}
        self.cfg = None

    def _on_log(self, level, msg, _user):
        text = msg.decode(errors="replace")
        tgt = self._log_target
        if callable(tgt) and not hasattr(tgt, "write"):
            tgt(level, text)
        else:
            tgt.write(text + "\n")

    # -- lifecycle --------------------------------------------------------

    def initialise_accelerator(self, input_sz, mode=PerSet, sync_pipeline=False, *, n_templates=85, max_taps=421,
                               tile=2048, n_harmonics=8, cands_per_harmonic=512, spill_per_harmonic=0, device=0,
                               reference_fft_order=False) -> bool:
        """FDAS::initialise_accelerator (FDAS.h:45-48).

        The bitstream / platform / device selectors of the reference have no
        meaning on a GPU; ``mode`` (DDR bank placement) is accepted and ignored.
        The compile-time GenInfo constants (gen_info.h) are keyword arguments.
        """
        self.close()
        cfg = N.FdasConfig()
        self._lib.fdas_default_config(C.byref(cfg))
        cfg.n_bins = int(input_sz)
        cfg.n_templates = int(n_templates)
        cfg.max_taps = int(max_taps)
        cfg.tile = int(tile)
        cfg.n_harmonics = int(n_harmonics)
        cfg.cands_per_harmonic = int(cands_per_harmonic)
        cfg.spill_per_harmonic = int(spill_per_harmonic)
        cfg.device = int(device)
        cfg.flags = (N.FLAG_REFERENCE_FFT_ORDER if reference_fft_order else 0) | \
                    (N.FLAG_SYNC_PIPELINE if sync_pipeline else 0)
        st = self._lib.fdas_create(C.byref(cfg), self._log_cb, None, C.byref(self._ctx))
        self._status = st
        if st != N.FDAS_OK:
            self._ctx = C.c_void_p()
            return False
        self.cfg = cfg
        return True

    def close(self):
        if self._ctx:
            self._lib.fdas_destroy(self._ctx)
            self._ctx = C.c_void_p()
        self._keep = {0: (), 1: ()}

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def last_error(self) -> str:
        if not self._ctx:
            return ""
        return self._lib.fdas_last_error(self._ctx).decode(errors="replace")

    @property
    def last_status(self) -> int:
        return self._status

    def _ok(self, st) -> bool:
        self._status = st
        return st == N.FDAS_OK

    # -- sizes (FDAS.h:78-88) ----------------------------------------------

    def _size(self, which):
        return int(self._lib.fdas_get_size(self._ctx, which))

    def get_input_sz(self):
        return self._size(N.SIZE_INPUT)

    def get_tiles_sz(self):
        return self._size(N.SIZE_TILES)

    def get_templates_sz(self):
        return self._size(N.SIZE_TEMPLATES)

    def get_thresholds_sz(self):
        return self._size(N.SIZE_THRESHOLDS)

    def get_fop_sz(self):
        return self._size(N.SIZE_FOP)

    def get_candidate_list_sz(self):
        return self._size(N.SIZE_CANDIDATES)

    # -- uploads -----------------------------------------------------------

    def upload_templates(self, templates, ab) -> bool:
        """FDAS::upload_templates (FDAS.h:50): T x S complex spectra."""
        t = _c64(templates, (self.get_templates_sz(),))
        return self._ok(self._lib.fdas_upload_templates(self._ctx, N.fptr(t.view(np.float32)), ab))

    def upload_template_taps(self, taps, ab) -> bool:
        """Time-domain taps (T x L complex); the library transforms them to the tile size."""
        taps = np.ascontiguousarray(taps, dtype=np.complex64)
        if taps.ndim != 2 or taps.shape[0] != self.cfg.n_templates:
            raise ValueError("taps must be (n_templates, n_taps)")
        return self._ok(self._lib.fdas_upload_template_taps(self._ctx, N.fptr(taps.view(np.float32)),
                                                            taps.shape[1], ab))

    def upload_thresholds(self, thresholds, ab) -> bool:
        thr = np.ascontiguousarray(thresholds, dtype=np.float32)
        return self._ok(self._lib.fdas_upload_thresholds(self._ctx, N.fptr(thr), ab))

    def perform_input_tiling(self, input, ab) -> bool:
        x = _c64(input, (self.get_input_sz(),))
        return self._ok(self._lib.fdas_perform_input_tiling(self._ctx, N.fptr(x.view(np.float32)), ab))

    def set_input_device(self, device_ptr: int, ab) -> bool:
        """Spectrum already in device memory (e.g. a torch tensor's data_ptr())."""
        return self._ok(self._lib.fdas_set_input_device(self._ctx, C.c_void_p(int(device_ptr)), ab))

    # -- processing --------------------------------------------------------

    def perform_ft_convolution(self, which, ab) -> bool:
        return self._ok(self._lib.fdas_perform_ft_convolution(self._ctx, which, ab))

    def perform_harmonic_summing(self, thresholds, which, ab) -> bool:
        if thresholds is None:
            return self._ok(self._lib.fdas_perform_harmonic_summing(self._ctx, None, which, ab))
        thr = np.ascontiguousarray(thresholds, dtype=np.float32)
        return self._ok(self._lib.fdas_perform_harmonic_summing(self._ctx, N.fptr(thr), which, ab))

    def _check_lists(self, detection_location, detection_power):
        """The library writes candidate_list_sz entries from each array's first element."""
        n = self.get_candidate_list_sz()
        if detection_location.dtype != np.uint32 or detection_location.size < n or \
                not detection_location.flags.c_contiguous:
            raise ValueError("detection_location must be a contiguous uint32 array of candidate_list_sz")
        if detection_power.dtype != np.float32 or detection_power.size < n or not detection_power.flags.c_contiguous:
            raise ValueError("detection_power must be a contiguous float32 array of candidate_list_sz")

    def launch(self, input, thresholds, detection_location, detection_power, which, ab) -> bool:
        """FDAS::launch (FDAS.h:60).  The arrays are kept alive until wait(ab)."""
        x = _c64(input, (self.get_input_sz(),))
        thr = np.ascontiguousarray(thresholds, dtype=np.float32)
        self._check_lists(detection_location, detection_power)
        self._keep[ab] = (x, thr, detection_location, detection_power)
        return self._ok(self._lib.fdas_launch(self._ctx, N.fptr(x.view(np.float32)), N.fptr(thr),
                                              N.uptr(detection_location), N.fptr(detection_power), which, ab))

    def launch_resident(self, thresholds, detection_location, detection_power, which, ab) -> bool:
        """launch() on the spectrum already resident in set ab (no host-to-device copy)."""
        thr = np.ascontiguousarray(thresholds, dtype=np.float32)
        self._check_lists(detection_location, detection_power)
        self._keep[ab] = (thr, detection_location, detection_power)
        return self._ok(self._lib.fdas_launch_resident(self._ctx, N.fptr(thr), N.uptr(detection_location),
                                                       N.fptr(detection_power), which, ab))

    def wait(self, ab) -> bool:
        ok = self._ok(self._lib.fdas_wait(self._ctx, ab))
        self._keep[ab] = ()
        return ok

    def end_pipeline(self, ab) -> bool:
        return self._ok(self._lib.fdas_end_pipeline(self._ctx, ab))

    # -- retrieval ---------------------------------------------------------

    def retrieve_tiles(self, tiles, ab) -> bool:
        if tiles.dtype != np.complex64 or tiles.size != self.get_tiles_sz() or not tiles.flags.c_contiguous:
            raise ValueError("tiles must be a contiguous complex64 array of tiles_sz")
        return self._ok(self._lib.fdas_retrieve_tiles(self._ctx, N.fptr(tiles.view(np.float32)), ab))

    def retrieve_FOP(self, fop, ab) -> bool:
        if fop.dtype != np.float32 or fop.size != self.get_fop_sz() or not fop.flags.c_contiguous:
            raise ValueError("fop must be a contiguous float32 array of fop_sz")
        return self._ok(self._lib.fdas_retrieve_fop(self._ctx, N.fptr(fop), ab))

    def inject_FOP(self, fop, ab) -> bool:
        f = np.ascontiguousarray(fop, dtype=np.float32)
        if f.size != self.get_fop_sz():
            raise ValueError("fop size mismatch")
        self._keep_fop = f
        return self._ok(self._lib.fdas_inject_fop(self._ctx, N.fptr(f), ab))

    def retrieve_candidates(self, detection_location, detection_power, ab) -> bool:
        self._check_lists(detection_location, detection_power)
        return self._ok(self._lib.fdas_retrieve_candidates(self._ctx, N.uptr(detection_location),
                                                           N.fptr(detection_power), ab))

    def retrieve_candidates_wide(self, ab):
        """Returns (harmonic, template_num, bin, power, n_detected) arrays; raises on failure."""
        n = self.get_candidate_list_sz()
        buf = (N.FdasCandidate * max(n, 1))()
        written = C.c_uint32(0)
        det = np.zeros(self.cfg.n_harmonics, dtype=np.uint32)
        st = self._lib.fdas_retrieve_candidates_wide(self._ctx, buf, n, C.byref(written), N.uptr(det), ab)
        self._status = st
        if st not in (N.FDAS_OK, 7):
            raise RuntimeError(self.last_error)
        arr = np.ctypeslib.as_array(buf)[:written.value]
        return (arr["harmonic"].astype(np.int64), arr["template_num"].astype(np.int64),
                arr["bin"].astype(np.int64), arr["power"].astype(np.float32), det)

    def retrieve_harmonic_plane(self, k, which, ab) -> np.ndarray:
        out = np.zeros(self.get_fop_sz(), dtype=np.float32)
        if not self._ok(self._lib.fdas_retrieve_harmonic_plane(self._ctx, int(k), N.fptr(out), which, ab)):
            raise RuntimeError(self.last_error)
        return out.reshape(self.cfg.n_templates, self.get_input_sz())

    # -- profiling (FDAS.h:74-76) -----------------------------------------

    def quantile_thresholds(self, n_per_plane, which, ab):
        """The reference's mock thresholds (hsum.py:125-131) from the FOP in set ab: about
        n_per_plane candidates per plane.  Returns float32[H], or None on failure."""
        thr = np.zeros(self.cfg.n_harmonics, np.float32)
        if not self._ok(self._lib.fdas_quantile_thresholds(self._ctx, int(n_per_plane), which, ab, N.fptr(thr))):
            return None
        return thr

    def print_stats(self, ab, reset=False):
        self._lib.fdas_print_stats(self._ctx, ab, 1 if reset else 0)

    def print_events(self, ab):
        self._lib.fdas_print_events(self._ctx, ab)

    def stage_times(self, ab) -> np.ndarray:
        """[input, FT-conv kernel, HSUM kernel, selection, read-back, total] in ms of the last run on ab."""
        out = np.zeros(6, dtype=np.float32)
        self._lib.fdas_get_stage_times(self._ctx, ab, N.fptr(out))
        return out


def make_drift_bank(drifts, max_taps: int) -> np.ndarray:
    """(n, max_taps) complex64 bank of drifting-tone responses (tmpl-gen.py:45-78, 154-159)."""
    lib = N.load()
    d = np.ascontiguousarray(drifts, dtype=np.float64)
    out = np.zeros((d.size, int(max_taps)), dtype=np.complex64)
    st = lib.fdas_make_drift_bank(N.dptr(d), d.size, int(max_taps), N.fptr(out.view(np.float32)))
    if st != N.FDAS_OK:
        raise ValueError(f"fdas_make_drift_bank failed: {N.STATUS_NAMES.get(st, st)}")
    return out


def reference_drifts(n_templates=85, max_accel=350.0, unaccel_freq=500.0, t_obs=600.0,
                     speed_of_light=299792458.0) -> np.ndarray:
    """Drifts (bins) of the reference generator's default bank (tmpl-gen.py:52-56)."""
    lib = N.load()
    out = np.zeros(int(n_templates), dtype=np.float64)
    st = lib.fdas_reference_drifts(int(n_templates), float(max_accel), float(unaccel_freq), float(t_obs),
                                   float(speed_of_light), N.dptr(out))
    if st != N.FDAS_OK:
        raise ValueError("fdas_reference_drifts failed")
    return out
