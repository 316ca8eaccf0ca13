import os
import sys

import numpy as np
import pytest
import torch  # noqa: F401  (first: its HIP runtime is the one the extensions then bind to)

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "demont-engine_amd", "python"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X); parity tests through the C ABI")


@pytest.fixture(scope="session")
def device():
    import dmt
    dev = dmt.Device(0)
    yield dev
    dev.close()   # deferred by the library while scenes are alive


def rel_mse(a: np.ndarray, b: np.ndarray) -> float:
    """sum((a - b)^2) / sum(b^2) over all pixels and channels."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = float((b * b).sum())
    return float(((a - b) ** 2).sum()) / max(den, 1e-30)
