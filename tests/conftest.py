"""Shared test setup: import paths for the product bindings and the oracle,
and the ``gpu`` marker (tests that need an MI355X)."""
import os
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(REPO, "gauss-legendre-spherical-t_amd")
for p in (PKG, REPO):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a GPU (MI355X) to run")


def gpu_available() -> bool:
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def cuda():
    if not gpu_available():
        pytest.skip("no GPU")
    import torch
    return torch.device("cuda:0")


def rel_err(a, b):
    """||a - b|| / ||b|| over all components (lists of arrays)."""
    import numpy as np
    num = sum(float(((np.asarray(x) - np.asarray(y)) ** 2).sum())
              for x, y in zip(a, b))
    den = sum(float((np.asarray(y) ** 2).sum()) for y in b)
    return (num / den) ** 0.5
