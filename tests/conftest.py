"""Shared fixtures.  GPU tests are marked ``gpu``; everything else runs on the CPU."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libcvchip on cuda:0)")


def have_gpu() -> bool:
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def orc():
    import oracle
    if not os.path.exists(os.path.join(ROOT, "oracle", "libcvcoracle.so")):
        oracle.build()
    return oracle


@pytest.fixture(scope="session")
def chip():
    """The product library; GPU tests fail loudly when it is missing."""
    import cvc_amd
    cvc_amd.lib()
    return cvc_amd


def rand_spinor(rng, V, n=None):
    shape = (V, 24) if n is None else (n, V, 24)
    return rng.standard_normal(shape)


def rel_err(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))
