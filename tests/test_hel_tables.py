"""CPU checks of the Helmholtz operator tables the library builds on the
host (wbfmm_tables.c): the coaxial (E|F) tables against the closed form
of their m = 0, n = 0 column, against the oracle's independent quad
precision tables, and through a rotate / S|R / rotate-back translation
built from the library's tables that must reproduce the direct field."""

import ctypes
import math

import numpy as np
import pytest
from scipy.special import spherical_jn, spherical_yn


def lib_table(wb, kind, L, kr):
    ne = (L + 1) * (L + 2) * (L + 3) // 6
    E = np.zeros(2 * ne)
    assert wb.lib().wbfmm_hip_hel_coaxial_table(kind, L, ctypes.c_double(kr), wb.ptr(E)) == 0
    return E


def cidx(l, m, n):
    return l * (l + 1) * (l + 2) // 6 + n * (n + 1) // 2 + m


@pytest.mark.parametrize("kind,kr", [(0, 0.4), (0, 3.7), (1, 0.9), (1, 6.2), (1, -2.5), (0, -1.3)])
def test_coaxial_column_closed_form(wb, kind, kr):
    # (E|F)_{l0}^0 = sqrt(2l+1) f_l(kr), f = j (R|R) or h = j + i y (S|R)
    # (G&D 2004 eq 4.83); f_l(-x) = (-1)^l f_l(x)
    L = 14
    E = lib_table(wb, kind, L, kr)
    x, sg = abs(kr), (-1) ** np.arange(L + 1) if kr < 0 else np.ones(L + 1)
    want = np.sqrt(2 * np.arange(L + 1) + 1) * sg * (
        spherical_jn(np.arange(L + 1), x) + (1j * spherical_yn(np.arange(L + 1), x) if kind else 0))
    got = np.array([E[2 * cidx(l, 0, 0)] + 1j * E[2 * cidx(l, 0, 0) + 1] for l in range(L + 1)])
    assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()


@pytest.mark.parametrize("kind,L,kr", [(0, 11, 0.87), (1, 11, 2.3), (1, 20, 9.5), (0, 25, 6.9)])
def test_coaxial_tables_match_oracle(wb, oracle, kind, L, kr):
    E = lib_table(wb, kind, L, kr)
    Eo = oracle.hel_coaxial_table(kind, L, kr)[:len(E)]
    assert np.abs(E - Eo).max() <= 1e-12 * np.abs(Eo).max()


def test_translation_from_library_tables_matches_direct(wb, oracle):
    # rotate -> S|R (library table) -> rotate back, evaluated at a target-box
    # point, reproduces the direct Helmholtz field (src/pass.c:73-102)
    k, N, w = 1.9, 20, 0.5
    off = (-3, 2, 2)
    rng = np.random.default_rng(7)
    src = rng.uniform(-0.5, 0.5, (6, 3)) * w
    q = rng.normal(size=6) + 1j * rng.normal(size=6)
    M = np.zeros(2 * (N + 1) ** 2)
    for s, qq in zip(src, q):
        oracle.hel_p2m(k, N, np.zeros(3), s, qq, M)
    th, ph, ch, r = oracle.shift_angles(*off)
    H = oracle.H_rotation(N, th)
    E = lib_table(wb, 1, N, k * r * w)
    Lc = oracle.hel_rotate(oracle.hel_coaxial(oracle.hel_rotate(M, N, H, ph, ch), N, N, E, N, 1),
                           N, H, ch, ph)
    xt0 = -np.array(off, float) * w
    xf = xt0 + rng.uniform(-0.5, 0.5, 3) * w
    fl = oracle.hel_j_eval(k, N, xt0, Lc, xf)
    fd = oracle.hel_direct(k, src, q, xf[None])[0]
    assert np.abs(fl - fd).max() < 1e-8 * np.abs(fd).max()
