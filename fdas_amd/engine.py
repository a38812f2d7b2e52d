"""Batched and served FDAS over the engine / plan entry points of ``include/fdas.h``.

``Engine`` holds template banks and a device arena bounded by an HBM budget.
``Engine.plan`` runs a fixed list of trials (uneven N, banks, harmonic counts)
in groups whose FT convolution overlaps the previous group's harmonic summing,
optionally replayed as a HIP graph; ``Engine.submit`` / ``Engine.poll`` serve
requests with continuous batching.  A trial's candidate lists equal those of
the single-trial path (``FDAS.launch`` + ``retrieve_candidates_wide``) at the
tile size the engine chose for its shape.
"""

from __future__ import annotations

import ctypes as C
import sys
import weakref
from dataclasses import dataclass, field

import numpy as np

from . import _native as N

KEY_BIN_BITS = 22
SCHEDULES = {"lagged": 0, "fused": 1, "streams": 2}  # fdas_engine_config.plan_schedule


class EngineError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(f"{N.STATUS_NAMES.get(status, status)}: {message}")
        self.status = status


@dataclass
class Trial:
    """One search request.  ``spectrum`` is a complex64 array (host) or a torch
    tensor on the GPU (complex64, or float32 (..., 2)); thresholds are absolute."""

    spectrum: object
    bank: int
    thresholds: np.ndarray
    part: int = 2
    tag: int = 0

    @property
    def n_harmonics(self):
        return len(self.thresholds)


@dataclass
class TrialResult:
    tag: int
    status: int
    tile: int
    n_detected: np.ndarray          # (H,) detections per plane before the cap
    n_kept: np.ndarray              # (H,)
    keys: list = field(default_factory=list)    # per plane: uint32 (row << 22 | bin), sorted
    powers: list = field(default_factory=list)  # per plane: float32
    latency_ms: float = 0.0

    def candidates(self, n_templates):
        """(harmonic, template t, bin, power) arrays over all planes."""
        c = n_templates // 2
        h, t, b, p = [], [], [], []
        for k, (ky, pw) in enumerate(zip(self.keys, self.powers)):
            h.append(np.full(len(ky), k + 1, np.uint32))
            t.append((ky >> KEY_BIN_BITS).astype(np.int32) - c)
            b.append(ky & ((1 << KEY_BIN_BITS) - 1))
            p.append(pw)
        cat = lambda xs, dt: np.concatenate(xs).astype(dt) if xs else np.zeros(0, dt)
        return cat(h, np.uint32), cat(t, np.int32), cat(b, np.uint32), cat(p, np.float32)


def _spectrum_ptr(spec):
    """(pointer, n_bins, flags, keep-alive) of a host array or a device tensor."""
    if hasattr(spec, "is_cuda"):
        import torch
        if not spec.is_cuda:
            spec = spec.numpy()
        else:
            if spec.dtype == torch.complex64:
                n = spec.numel()
            elif spec.dtype == torch.float32 and spec.shape[-1] == 2:
                n = spec.numel() // 2
            else:
                raise TypeError("device spectrum must be complex64 or float32 (..., 2)")
            if not spec.is_contiguous():
                raise ValueError("device spectrum must be contiguous")
            return spec.data_ptr(), n, N.REQ_DEVICE_SPECTRUM, spec
    a = np.ascontiguousarray(spec, dtype=np.complex64).reshape(-1)
    return a.ctypes.data, a.size, 0, a


class _Outputs:
    """Caller-owned list buffers of one trial."""

    def __init__(self, n_harmonics, cap):
        self.key = np.zeros(n_harmonics * cap, np.uint32)
        self.power = np.zeros(n_harmonics * cap, np.float32)
        self.res = N.Result()
        self.res.key = self.key.ctypes.data_as(C.POINTER(C.c_uint32))
        self.res.power = self.power.ctypes.data_as(C.POINTER(C.c_float))

    def result(self, cap):
        r = self.res
        H = r.n_harmonics
        det = np.array(r.n_detected[:H], np.uint32)
        kept = np.array(r.n_kept[:H], np.uint32)
        keys = [self.key[k * cap:k * cap + kept[k]].copy() for k in range(H)]
        pw = [self.power[k * cap:k * cap + kept[k]].copy() for k in range(H)]
        return TrialResult(int(r.tag), int(r.status), int(r.tile), det, kept, keys, pw, float(r.latency_ms))


def _request(trial, keep):
    ptr, n, flags, alive = _spectrum_ptr(trial.spectrum)
    keep.append(alive)
    thr = np.asarray(trial.thresholds, np.float32).reshape(-1)
    if not 1 <= thr.size <= N.MAX_HARMONICS:
        raise ValueError(f"1..{N.MAX_HARMONICS} thresholds expected, got {thr.size}")
    q = N.Request()
    q.spectrum = ptr
    q.n_bins = n
    q.bank = trial.bank
    q.n_harmonics = thr.size
    q.part = trial.part
    q.flags = flags
    for k in range(thr.size):
        q.thresholds[k] = float(thr[k])
    q.tag = trial.tag
    return q


class Plan:
    def __init__(self, engine, trials, graph):
        self._eng = engine
        self._lib = engine._lib
        self._keep = []
        reqs = (N.Request * len(trials))(*[_request(t, self._keep) for t in trials])
        self._reqs = reqs
        self._h = C.c_void_p()
        engine._check(self._lib.fdas_plan_create(engine._h, reqs, len(trials), N.PLAN_GRAPH if graph else 0,
                                                 C.byref(self._h)))
        engine._plans.add(self)
        self.n = len(trials)
        self._out = [_Outputs(t.n_harmonics, engine.cap) for t in trials]
        self._res = (N.Result * self.n)()

    def set_trial(self, i, spectrum=None, thresholds=None):
        ptr = None
        if spectrum is not None:
            p, n, flags, alive = _spectrum_ptr(spectrum)
            if n != self._reqs[i].n_bins or flags != self._reqs[i].flags:
                raise ValueError("a replacement spectrum must have the same length and memory kind")
            self._keep.append(alive)
            ptr = p
        thr = None
        if thresholds is not None:
            thr = np.ascontiguousarray(thresholds, np.float32)
            self._keep.append(thr)
        self._eng._check(self._lib.fdas_plan_set_request(self._h, i, ptr, N.fptr(thr) if thr is not None else None))

    def launch(self):
        self._eng._check(self._lib.fdas_plan_launch(self._h))

    def wait(self, results=True, raise_overflow=True):
        if not results:
            self._eng._check(self._lib.fdas_plan_wait(self._h, None))
            return None
        for i, o in enumerate(self._out):
            self._res[i].key = o.res.key
            self._res[i].power = o.res.power
        st = self._lib.fdas_plan_wait(self._h, self._res)
        if st not in (N.FDAS_OK, 7) or (st == 7 and raise_overflow):
            self._eng._check(st)
        out = []
        for i, o in enumerate(self._out):
            C.memmove(C.byref(o.res), C.byref(self._res[i]), C.sizeof(N.Result))
            out.append(o.result(self._eng.cap))
        return out

    def run(self):
        self.launch()
        return self.wait()

    def info(self):
        i = N.PlanInfo()
        self._eng._check(self._lib.fdas_plan_get_info(self._h, C.byref(i)))
        return {f: getattr(i, f) for f, _ in N.PlanInfo._fields_}

    def close(self):
        if self._h:
            self._lib.fdas_plan_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Engine:
    def __init__(self, *, tile=0, rows_per_item=0, hsum_rows=0, cands_per_harmonic=4096, spill_per_harmonic=0,
                 group_trials=8, groups_in_flight=3, device=0, hbm_budget_bytes=200 * 10**9,
                 ftc_workgroups_per_cu=0, schedule="lagged", log=None):
        self._lib = N.load()
        self._log_target = log
        self._log_cb = N.LOG_FN(self._on_log)
        cfg = N.EngineConfig()
        self._lib.fdas_engine_default_config(C.byref(cfg))
        cfg.tile, cfg.rows_per_item, cfg.hsum_rows = tile, rows_per_item, hsum_rows
        cfg.cands_per_harmonic, cfg.spill_per_harmonic = cands_per_harmonic, spill_per_harmonic
        cfg.group_trials, cfg.groups_in_flight, cfg.device = group_trials, groups_in_flight, device
        cfg.hbm_budget_bytes = hbm_budget_bytes
        cfg.ftc_workgroups_per_cu = ftc_workgroups_per_cu
        cfg.plan_schedule = SCHEDULES[schedule]
        self._h = C.c_void_p()
        self._plans = weakref.WeakSet()  # closed before the engine: fdas_engine_destroy frees its plans
        self.last_error = ""
        st = self._lib.fdas_engine_create(C.byref(cfg), self._log_cb, None, C.byref(self._h))
        if st != N.FDAS_OK:
            raise EngineError(st, self.last_error)
        self.cap = cands_per_harmonic
        self.banks = {}
        self._pending = {}

    def _on_log(self, level, msg, _user):
        text = msg.decode(errors="replace")
        if level == 2:
            self.last_error = text
        tgt = self._log_target
        if tgt is None:
            if level >= 1:
                print(text, file=sys.stderr)
        elif callable(tgt) and not hasattr(tgt, "write"):
            tgt(level, text)
        else:
            tgt.write(text + "\n")

    def _check(self, st):
        if st != N.FDAS_OK:
            msg = self._lib.fdas_engine_last_error(self._h).decode(errors="replace") if self._h else self.last_error
            raise EngineError(st, msg)

    def add_bank(self, taps):
        taps = np.ascontiguousarray(taps, np.complex64)
        if taps.ndim != 2:
            raise ValueError("taps must be (n_templates, n_taps)")
        bid = C.c_uint32()
        self._check(self._lib.fdas_engine_add_bank(self._h, taps.ctypes.data_as(N._F), taps.shape[0], taps.shape[1],
                                                   C.byref(bid)))
        self.banks[bid.value] = taps.shape[0]
        return bid.value

    def prepare(self, n_bins, bank, part=2):
        tile = C.c_uint32()
        self._check(self._lib.fdas_engine_prepare(self._h, n_bins, bank, part, C.byref(tile)))
        return tile.value

    def shape_info(self, n_bins, bank, part=2):
        """The work shape fixed for (n_bins, bank, part) and its per-trial transform counts."""
        si = N.ShapeInfo()
        self._check(self._lib.fdas_engine_shape_info(self._h, n_bins, bank, part, C.byref(si)))
        return {f: int(getattr(si, f)) for f, _ in N.ShapeInfo._fields_}

    def tune_cache(self, path):
        """Read tuned shapes from `path` and append every shape tuned later (fdas_engine_tune_cache)."""
        self._check(self._lib.fdas_engine_tune_cache(self._h, str(path).encode()))

    def plan(self, trials, graph=False):
        return Plan(self, trials, graph)

    def run(self, trials, graph=False):
        p = self.plan(trials, graph)
        try:
            return p.run()
        finally:
            p.close()

    # -- serving ------------------------------------------------------------

    def submit(self, trial):
        keep = []
        q = _request(trial, keep)
        out = _Outputs(trial.n_harmonics, self.cap)
        addr = C.addressof(out.res)
        self._pending[addr] = (out, keep)  # before submit: a poller thread may see it complete at once
        st = self._lib.fdas_engine_submit(self._h, C.byref(q), C.byref(out.res))
        if st != N.FDAS_OK:
            self._pending.pop(addr, None)
            self._check(st)
        return out

    def poll(self, max_results=64, timeout_ms=-1):
        arr = (C.POINTER(N.Result) * max_results)()
        n = C.c_uint32()
        self._check(self._lib.fdas_engine_poll(self._h, arr, max_results, C.byref(n), timeout_ms))
        done = []
        for i in range(n.value):
            addr = C.addressof(arr[i].contents)
            out, _ = self._pending.pop(addr)
            done.append(out.result(self.cap))
        return done

    def cancel(self, tag):
        """Withdraw queued requests with this tag (fdas_engine_cancel); returns how many.  They complete
        with status FDAS_ECANCELED and come back from poll()."""
        n = C.c_uint32()
        self._check(self._lib.fdas_engine_cancel(self._h, C.c_uint64(tag), C.byref(n)))
        return n.value

    def drain(self):
        out = []
        while self._pending:
            out.extend(self.poll(timeout_ms=-1))
        return out

    def stats(self):
        s = N.EngineStats()
        self._check(self._lib.fdas_engine_get_stats(self._h, C.byref(s)))
        return {f: getattr(s, f) for f, _ in N.EngineStats._fields_ if f != "pad_"}

    def close(self):
        if self._h:
            for p in list(self._plans):
                p.close()
            self._lib.fdas_engine_destroy(self._h)
            self._h = C.c_void_p()
            self._pending.clear()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
