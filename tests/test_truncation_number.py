"""Truncation numbers of Gumerov & Duraiswami (JASA 2009), the rule the
reference uses to pick Helmholtz orders per level (src/util.c:474-503).

Pinned against values computed independently of both the oracle and the
library (numpy, for a domain of width D = 2.02 at levels 2..7), and by
properties of the formula: the low-frequency limit is the separation-ratio
bound alone, and the order never decreases with k or with accuracy.
Runs on the CPU: the library's wbfmm_truncation_number only reads the
tree's width.
"""

import ctypes
import math

import pytest

# (k, tol) -> orders at levels 2..7 for D = 2.02
KNOWN = {
    (16.0, 1e-6): [25, 23, 23, 23, 23, 23],
    (16.0, 1e-8): [31, 30, 30, 30, 30, 30],
    (32.0, 1e-8): [37, 31, 30, 30, 30, 30],
    (20.0, 1e-4): [21, 17, 17, 16, 16, 16],
}


def product_tn(W, k, D, level, tol):
    t = W.WbfmmTree()
    t.D = D
    return W.lib().wbfmm_truncation_number(ctypes.byref(t), k, level, tol)


@pytest.mark.parametrize("key", sorted(KNOWN))
def test_oracle_truncation_number_known_values(oracle, key):
    k, tol = key
    got = [oracle.lib().orc_truncation_number(k, 2.02, l, tol) for l in range(2, 8)]
    assert got == KNOWN[key]


@pytest.mark.parametrize("key", sorted(KNOWN))
def test_library_truncation_number_known_values(wb, key):
    k, tol = key
    got = [product_tn(wb, k, 2.02, l, tol) for l in range(2, 8)]
    assert got == KNOWN[key]


def test_low_frequency_limit(wb, oracle):
    # as k -> 0 only the separation-ratio bound (ratio 2) remains:
    # p = ceil(1 - log2(tol (1 - 1/2) sqrt(1 - 1/2)))
    for tol in (1e-3, 1e-6, 1e-10):
        want = math.ceil(1.0 - math.log2(tol * 0.5 * math.sqrt(0.5)))
        assert oracle.lib().orc_truncation_number(1e-12, 1.0, 3, tol) == want
        assert product_tn(wb, 1e-12, 1.0, 3, tol) == want


def test_monotone_in_k_and_accuracy(wb, oracle):
    for level in range(2, 8):
        prev = 0
        for k in (0.5, 1, 2, 4, 8, 16, 32, 64):
            p = product_tn(wb, k, 2.0, level, 1e-6)
            assert p >= prev
            assert p == oracle.lib().orc_truncation_number(k, 2.0, level, 1e-6)
            prev = p
        prev = 0
        for tol in (1e-2, 1e-4, 1e-6, 1e-8, 1e-12):
            p = product_tn(wb, 10.0, 2.0, level, tol)
            assert p >= prev
            prev = p
