"""Front end of the search: SIGPROC time series in, spectrum out.

``read_tim`` follows the reference's reader (``ft_conv.py:39-67``) through the
library's C implementation; ``FrontEnd`` runs the spectrum transform
X[f] = DFT(s)[f] / Ns, f < N, Ns = 2N (``ft_conv.py:123-146``) on the GPU in
FP32 through ``fdas_frontend_*``.
"""

from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as N


class FrontEndError(RuntimeError):
    pass


def read_tim(path, read_header=True):
    """(float32 samples, t_samp or None) of a SIGPROC time-series file."""
    lib = N.load()
    n = C.c_uint64()
    ts = C.c_double()
    st = lib.fdas_read_tim(str(path).encode(), int(read_header), None, 0, C.byref(n), C.byref(ts))
    if st != N.FDAS_OK:
        raise FrontEndError(lib.fdas_io_last_error().decode())
    out = np.empty(n.value, np.float32)
    st = lib.fdas_read_tim(str(path).encode(), int(read_header), N.fptr(out), out.size, C.byref(n), C.byref(ts))
    if st != N.FDAS_OK:
        raise FrontEndError(lib.fdas_io_last_error().decode())
    return out, (ts.value if ts.value != 0.0 else None)


def trim_series(samples, n_bins):
    """The first 2N samples, as the reference keeps them (ft_conv.py:123-133)."""
    if samples.size < 2 * n_bins:
        raise ValueError(f"input contains {samples.size} samples, but {2 * n_bins} are needed")
    return np.ascontiguousarray(samples[:2 * n_bins], np.float32)


class FrontEnd:
    """Device spectrum transform for N bins (a power of two, 2 .. 2^22)."""

    def __init__(self, n_bins, device=0):
        self._lib = N.load()
        self.n_bins = int(n_bins)
        self._h = C.c_void_p()
        st = self._lib.fdas_frontend_create(self.n_bins, device, C.byref(self._h))
        if st != N.FDAS_OK:
            raise FrontEndError(f"{N.STATUS_NAMES.get(st, st)}: cannot create a front end for N={n_bins} "
                                f"(a power of two, 2 .. 2^22, on a GPU)")

    def run(self, series, out=None):
        """Spectrum of 2N samples.  ``series``: a float32 array or a torch tensor on
        the GPU; ``out``: None (returns a complex64 array) or a complex64 torch
        tensor on the GPU that receives the spectrum."""
        flags = 0
        keep = []
        if hasattr(series, "is_cuda") and series.is_cuda:
            if series.numel() != 2 * self.n_bins or not series.is_contiguous():
                raise ValueError("series must hold 2N contiguous float32 samples")
            sptr = series.data_ptr()
            flags |= N.FE_DEVICE_INPUT
        else:
            s = np.ascontiguousarray(series, np.float32).reshape(-1)
            if s.size != 2 * self.n_bins:
                raise ValueError(f"series must hold {2 * self.n_bins} samples, got {s.size}")
            keep.append(s)
            sptr = s.ctypes.data
        if out is None:
            res = np.empty(self.n_bins, np.complex64)
            optr = res.ctypes.data
        else:
            if not (hasattr(out, "is_cuda") and out.is_cuda) or out.numel() != self.n_bins:
                raise ValueError("out must be a complex64 GPU tensor of N elements")
            res = out
            optr = out.data_ptr()
            flags |= N.FE_DEVICE_OUTPUT
        st = self._lib.fdas_frontend_run(self._h, C.c_void_p(sptr), flags, C.c_void_p(optr))
        if st != N.FDAS_OK:
            raise FrontEndError(self._lib.fdas_frontend_last_error(self._h).decode())
        return res

    def close(self):
        if self._h:
            self._lib.fdas_frontend_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
