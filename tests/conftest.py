import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); runs the HIP path through the C ABI")


def _ensure_built():
    from application_classification_amd import _lib

    if not os.path.exists(_lib.LIB_PATH):
        from application_classification_amd import build

        build.build()
    from oracle import oracle

    if not os.path.exists(oracle.LIB_PATH):
        oracle.build()


_ensure_built()


def allclose_report(a, b, rtol=1e-5, atol=1e-8):
    """The reference's own comparison (validate.py:7, np.allclose defaults): |a - b| <= atol + rtol * |b|."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    err = np.abs(a - b)
    tol = atol + rtol * np.abs(b)
    ratio = float((err / tol).max()) if err.size else 0.0
    bad = int((err > tol).sum())
    return bad, ratio


@pytest.fixture(scope="session")
def gpu_available():
    import torch

    return torch.cuda.is_available()
