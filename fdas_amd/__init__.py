"""MI355X-native Fourier Domain Acceleration Search (FDAS) hot path.

The device path is ``libfdas_amd.so`` (HIP kernels for gfx950 behind the C ABI
in ``include/fdas.h``); this package is its Python face:

* :class:`fdas_amd.host.FDAS` mirrors the reference host class ``FDAS``;
* :mod:`fdas_amd.synth` generates the synthetic dedispersed time series and
  spectra used by the tests and the benchmark.
"""

from .host import FDAS, make_drift_bank, reference_drifts  # noqa: F401
from ._native import LIB_PATH, load  # noqa: F401

__all__ = ["FDAS", "make_drift_bank", "reference_drifts", "LIB_PATH", "load"]
