import os
import sys

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(REPO, "extended-rabitq_amd"))
sys.path.insert(0, REPO)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def small_data():
    """20k x 128 isotropic mixture, 64 k-means clusters, 64 queries."""
    from erq import data
    x, cen = data.mixture_isotropic(20000, 128, 64, 0.3, seed=1)
    q, _ = data.mixture_isotropic(64, 128, 64, 0.3, seed=2, centres=cen)
    c, assign = data.kmeans(x, 64, 6, seed=4)
    return x, q, c, assign


_index_cache = {}


def build_cached(key, fn):
    if key not in _index_cache:
        _index_cache[key] = fn()
    return _index_cache[key]


def rotate_np(idx, q):
    qp = np.zeros((q.shape[0], idx.D), np.float32)
    qp[:, :q.shape[1]] = q
    return (qp.astype(np.float64) @ idx.P.astype(np.float64)).astype(np.float32)
