import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


def gaussian_mixture(n, dim, clusters, std, seed, nq=0, qstd=0.05):
    """Synthetic base set (cluster centers N(0, 1), per-cluster std) and
    queries drawn as perturbed base points."""
    rng = np.random.default_rng(seed)
    centers = rng.normal(size=(clusters, dim)).astype(np.float32)
    x = (centers[rng.integers(0, clusters, n)] + std * rng.normal(size=(n, dim))).astype(np.float32)
    q = None
    if nq:
        q = (x[rng.integers(0, n, nq)] + qstd * rng.normal(size=(nq, dim))).astype(np.float32)
    return x, q


@pytest.fixture(scope="session")
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    torch.cuda.set_device(0)
    return torch


@pytest.fixture(scope="session")
def golden():
    """Vectors written by the reference bridge (tests/golden/make_golden.py)."""
    return np.load(os.path.join(ROOT, "tests", "golden", "ref_vectors.npz"))


@pytest.fixture(scope="session")
def golden_ss():
    """Vectors written by the reference's 1-bit query-codec scan
    (tests/golden/make_signscan_golden.py)."""
    return np.load(os.path.join(ROOT, "tests", "golden", "signscan_ref.npz"))
