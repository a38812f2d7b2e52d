import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

GOLDEN = ROOT / "tests" / "golden"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X); calls the library through the C ABI")
    config.addinivalue_line("markers", "slow: full-size (N = 2^21) parity case")


@pytest.fixture(scope="session")
def golden_small():
    with np.load(GOLDEN / "golden_small.npz", allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.fixture(scope="session")
def reference_bank():
    """fdas/tmpl/fdas_templates_85_350.npy of the reference (85 x 421 complex64)."""
    return np.load(GOLDEN / "fdas_templates_85_350.npy", allow_pickle=False)


@pytest.fixture(scope="session")
def reference_bank_ft_p4():
    """fdas/tmpl/fdas_templates_85_350_ft_p4.npy of the reference (85 x 2048 complex64, FFT order p=4)."""
    return np.load(GOLDEN / "fdas_templates_85_350_ft_p4.npy", allow_pickle=False)
