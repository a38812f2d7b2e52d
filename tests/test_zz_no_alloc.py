"""fdas_hsum_noalloc (the allocation-free summing entry point) against fdas_perform_harmonic_summing.

Both run the same kernels on the same FOP; the no-allocation entry point keeps its
scratch in a file-scope device array and writes the lists to the caller's device
memory.  Two calls in a row (the second smaller, with other T and H) must each give
the single-trial path's lists bit for bit.
"""

import ctypes as C

import numpy as np
import pytest

from fdas_amd import _native as N
from fdas_amd.host import FDAS
from oracle import fdas_oracle as O

from test_gpu_parity import make, wide_by_plane

pytestmark = pytest.mark.gpu


def _thresholds(fop, n_harmonics, n_per_plane):
    """The mock thresholds of hsum.py:125-131: about n_per_plane detections per plane."""
    return O.quantile_thresholds(O.harmonic_planes(fop, n_harmonics, dtype=np.float32), n_per_plane)


def _reference(fop, n_templates, n_harmonics, thr, cap, part):
    n = fop.shape[1]
    f = make(n, n_templates=n_templates, n_harmonics=n_harmonics, cands_per_harmonic=cap)
    assert f.inject_FOP(fop.reshape(-1), FDAS.A), f.last_error
    f.perform_harmonic_summing(thr, part, FDAS.A)
    planes, det = wide_by_plane(f, FDAS.A, n_harmonics)
    f.close()
    return planes, det


def _noalloc(torch, fop, n_templates, n_harmonics, thr, cap, part, spill):
    lib = N.load()
    d_fop = torch.from_numpy(fop).cuda()
    d_key = torch.zeros(n_harmonics * cap, dtype=torch.int32, device="cuda")
    d_pwr = torch.zeros(n_harmonics * cap, dtype=torch.float32, device="cuda")
    d_cnt = torch.full((2 * n_harmonics,), 7, dtype=torch.int32, device="cuda")
    msgs = []
    cb = N.LOG_FN(lambda lvl, m, u: msgs.append(m.decode()))
    thr = np.ascontiguousarray(thr, np.float32)
    torch.cuda.synchronize()
    st = lib.fdas_hsum_noalloc(C.c_void_p(d_fop.data_ptr()), fop.shape[1], n_templates, part, n_harmonics,
                               thr.ctypes.data_as(N._F), cap, spill, C.c_void_p(d_key.data_ptr()),
                               C.c_void_p(d_pwr.data_ptr()), C.c_void_p(d_cnt.data_ptr()),
                               C.c_void_p(torch.cuda.current_stream().cuda_stream), cb, None)
    assert st == N.FDAS_OK, msgs
    torch.cuda.synchronize()
    key = d_key.cpu().numpy().view(np.uint32)
    pwr = d_pwr.cpu().numpy()
    cnt = d_cnt.cpu().numpy().view(np.uint32)
    c = n_templates // 2
    planes = []
    for p in range(n_harmonics):
        m = int(cnt[p])
        k = key[p * cap:p * cap + m]
        planes.append(((k >> 22).astype(np.int64) - c, (k & 0x3FFFFF).astype(np.int64), pwr[p * cap:p * cap + m]))
    return planes, cnt[n_harmonics:]


def test_noalloc_equals_perform_harmonic_summing():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(91)
    # (templates, bins, harmonics, cap, part, spill per plane: 0 = the default rule, which needs more
    # than the fixed scratch at 16 planes)
    cases = [(85, 20011, 8, 256, FDAS.AllAccelerations, 0), (169, 5003, 16, 128, FDAS.NegativeAccelerations, 65536)]
    for T, n, H, cap, part, spill in cases:  # the second call is smaller than the first
        fop = rng.exponential(size=(T, n)).astype(np.float32)
        thr = _thresholds(fop, H, 600)
        ref_planes, ref_det = _reference(fop, T, H, thr, cap, part)
        planes, det = _noalloc(torch, fop, T, H, thr, cap, part, spill)
        assert np.array_equal(det, ref_det)
        assert any(int(d) > cap for d in det)  # the cap binds on some plane
        for (rt, rb, rp), (gt, gb, gp) in zip(ref_planes, planes):
            assert np.array_equal(rt, gt) and np.array_equal(rb, gb) and np.array_equal(rp, gp)
