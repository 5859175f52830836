import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


def splitmix64(seed, n):
    """Portable SplitMix64 stream -> n uint64 (same generator as bench.py)."""
    out = np.empty(n, dtype=np.uint64)
    s = np.uint64(seed)
    with np.errstate(over="ignore"):
        for i in range(n):
            s = s + np.uint64(0x9E3779B97F4A7C15)
            z = s
            z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
            z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
            out[i] = z ^ (z >> np.uint64(31))
    return out


def fit_box(*pts):
    """domain fitted as in tools/wbfmm-tree-calc.c.in:261-267: min corner
    minus 0.01, width plus 0.02"""
    allp = np.concatenate([p for p in pts if p is not None])
    lo = allp.min(0)
    hi = allp.max(0)
    D = float((hi - lo).max())
    return lo - 0.01, D + 0.02


@pytest.fixture(scope="session")
def oracle():
    import oracle as O
    O.build()
    return O


@pytest.fixture(scope="session")
def wb():
    import wbfmm_amd as W
    W.lib()
    return W
