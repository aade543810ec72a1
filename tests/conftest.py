import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    O.lib()
    return O


@pytest.fixture(scope="session")
def ccc():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import cccollisions_amd
    return cccollisions_amd


def sorted_rows(a):
    a = np.asarray(a)
    if a.size == 0:
        return a.reshape(0, a.shape[-1] if a.ndim > 1 else 1)
    a = a.reshape(-1, a.shape[-1])
    return a[np.lexsort(a.T[::-1])]


def valid_rows(a):
    """Rows of an output padded with -1 whose indices are all valid."""
    a = np.asarray(a)
    return a[(a >= 0).all(axis=-1)]
