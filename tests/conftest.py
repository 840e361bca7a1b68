"""Shared test set-up: the ``gpu`` marker, import paths and seeded inputs."""

from __future__ import annotations

import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU (runs the HIP kernels)")
    config.addinivalue_line("markers", "slow: long-running")


def gpu_available() -> bool:
    try:
        import torch

        return torch.cuda.is_available()
    except Exception:  # pragma: no cover
        return False


def f32(a):
    """Round to float32 once and return float64 (the inputs every path sees)."""
    return np.asarray(a, dtype=np.float32).astype(np.float64)


def speckle(seed: int, n: int, xlim=(-0.02, 0.02), zlim=(0.005, 0.045)):
    rng = np.random.default_rng(seed)
    x = rng.uniform(*xlim, n)
    z = rng.uniform(*zlim, n)
    rc = rng.standard_normal(n)
    return f32(x), f32(z), f32(rc)


def rel_l2(a, b) -> float:
    """||a - b|| / ||b|| in float64 (complex128 for complex inputs). An all-zero reference gives 0 when
    ``a`` is all zero too and inf otherwise, never NaN."""
    wide = np.complex128 if (np.iscomplexobj(a) or np.iscomplexobj(b)) else np.float64
    a = np.asarray(a, dtype=wide).ravel()
    b = np.asarray(b, dtype=wide).ravel()
    num = float(np.linalg.norm(a - b))
    den = float(np.linalg.norm(b))
    if den == 0.0:
        return 0.0 if num == 0.0 else float("inf")
    return num / den


@pytest.fixture(scope="session")
def cuda():
    if not gpu_available():
        pytest.skip("no ROCm GPU visible")
    import torch

    return torch.device("cuda", 0)
