"""Pins the CPU oracle's Helmholtz path (G = h0(kr)/(4 pi), complex
coefficients) against closed forms: scipy's spherical Bessel functions,
FP64 direct sums (and 50-digit mpmath sums in the golden fixture),
finite-difference gradients, and the FMM's convergence in the order."""

import json
import math
import os

import numpy as np
import pytest
from scipy.special import spherical_jn, spherical_yn

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def test_bessel_functions_match_scipy(oracle):
    import ctypes
    L = oracle.lib()
    for x in (0.05, 0.7, 3.0, 11.0, 30.0):
        N = 30
        j = np.zeros(N + 1)
        L.orc_bessel_j(N, x, oracle.P(j))
        want = spherical_jn(np.arange(N + 1), x)
        # the series branch below n+1 and the recursion above agree with
        # scipy; upward recursion loses a few digits near x ~ n (6e-11 at
        # x = n = 30), as in the reference's j_n (src/util.c:147-175)
        assert np.allclose(j, want, rtol=1e-9, atol=1e-300)
        h = np.zeros(2 * (N + 1))
        L.orc_bessel_h(N, x, oracle.P(h))
        # h_n by upward recursion (src/util.c:177-202): y_n is accurate for
        # every n, the j_n part only while n <= x (where the reference's
        # S|R tables use it: kr >= 2 k w at list-4 distances)
        n_ok = min(N, int(x)) + 1
        assert np.allclose(h[0:2 * n_ok:2], spherical_jn(np.arange(n_ok), x), rtol=1e-9, atol=1e-14)
        assert np.allclose(h[1::2], spherical_yn(np.arange(N + 1), x), rtol=1e-10)


def test_p2m_and_singular_evaluation(oracle):
    rng = np.random.default_rng(1)
    for k in (0.5, 2.0, 8.0):
        N = 24
        src = rng.uniform(-0.3, 0.3, (10, 3))
        q = rng.normal(size=10) + 1j * rng.normal(size=10)
        M = np.zeros(2 * (N + 1) ** 2)
        for s, qq in zip(src, q):
            oracle.hel_p2m(k, N, np.zeros(3), s, qq, M)
        xf = np.array([1.3, -0.9, 1.1])
        fd = oracle.hel_direct(k, src, q, xf[None])[0, 0]
        assert abs(oracle.hel_h_eval(k, N, np.zeros(3), M, xf) - fd) < 1e-10 * abs(fd)


def test_dipole_sign_and_expansion(oracle):
    # the Helmholtz dipole is d n.grad_x G = -k h1(kr) d n.(x-y)/(4 pi r)
    # (src/local.c:155-166), checked in closed form and through P2M
    k = 1.7
    y = np.array([0.1, -0.2, 0.05])
    n = np.array([0.0, 0.6, 0.8])
    x = np.array([0.9, 0.4, -0.7])
    d = 0.3 - 0.8j
    r = np.linalg.norm(x - y)
    h1 = spherical_jn(1, k * r) + 1j * spherical_yn(1, k * r)
    want = -k * h1 * d * (n @ (x - y)) / (4 * np.pi * r)
    got = oracle.hel_direct(k, y[None], None, x[None], nrm=n[None], d=np.array([d]))[0, 0]
    assert abs(got - want) < 1e-14 * abs(want)
    N = 22
    M = oracle.hel_p2m_normal(k, N, np.zeros(3), y, n, d)
    assert abs(oracle.hel_h_eval(k, N, np.zeros(3), M, x * 3) -
               oracle.hel_direct(k, y[None], None, (x * 3)[None], nrm=n[None], d=np.array([d]))[0, 0]) < 1e-10


def test_gradients_match_finite_differences(oracle):
    rng = np.random.default_rng(2)
    k, h = 2.0, 1e-6
    N = 10
    Lc = rng.normal(size=2 * (N + 1) ** 2)
    xf = np.array([0.2, -0.3, 0.25])
    g = oracle.hel_j_eval(k, N, np.zeros(3), Lc, xf)[1:]
    fd = np.array([(oracle.hel_j_eval(k, N, np.zeros(3), Lc, xf + h * e, 1)[0]
                    - oracle.hel_j_eval(k, N, np.zeros(3), Lc, xf - h * e, 1)[0]) / (2 * h)
                   for e in np.eye(3)])
    assert np.abs(g - fd).max() < 1e-7 * np.abs(g).max()
    src = rng.uniform(-1, 1, (6, 3))
    q = rng.normal(size=6) + 1j * rng.normal(size=6)
    nrm = rng.normal(size=(6, 3))
    nrm /= np.linalg.norm(nrm, axis=1)[:, None]
    d = rng.normal(size=6) + 1j * rng.normal(size=6)
    xf = np.array([1.5, 0.3, -1.2])
    g = oracle.hel_direct(k, src, q, xf[None], nrm=nrm, d=d)[0, 1:]
    fd = np.array([(oracle.hel_direct(k, src, q, (xf + h * e)[None], nrm=nrm, d=d)[0, 0]
                    - oracle.hel_direct(k, src, q, (xf - h * e)[None], nrm=nrm, d=d)[0, 0]) / (2 * h)
                   for e in np.eye(3)])
    assert np.abs(g - fd).max() < 1e-7 * np.abs(g).max()


@pytest.mark.parametrize("off", [(2, 0, 0), (-3, 1, 2), (0, 0, 2), (0, 0, -3), (3, -3, 3)])
def test_m2l_rotate_coaxial_rotate(oracle, off):
    # rotate -> S|R -> rotate back (src/pass.c:73-102)
    rng = np.random.default_rng(sum(off) + 10)
    k, N, w = 2.0, 22, 0.5
    xs0 = np.zeros(3)
    xt0 = xs0 - np.array(off, float) * w
    src = xs0 + rng.uniform(-0.5, 0.5, (8, 3)) * w
    q = rng.normal(size=8) + 1j * rng.normal(size=8)
    M = np.zeros(2 * (N + 1) ** 2)
    for s, qq in zip(src, q):
        oracle.hel_p2m(k, N, xs0, s, qq, M)
    th, ph, ch, r = oracle.shift_angles(*off)
    H = oracle.H_rotation(N, th)
    E = oracle.hel_coaxial_table(1, N, k * r * w)
    Lc = oracle.hel_rotate(oracle.hel_coaxial(oracle.hel_rotate(M, N, H, ph, ch), N, N, E, N, 1),
                           N, H, ch, ph)
    xf = xt0 + rng.uniform(-0.5, 0.5, 3) * w
    fl = oracle.hel_j_eval(k, N, xt0, Lc, xf)
    fd = oracle.hel_direct(k, src, q, xf[None])[0]
    assert np.abs(fl - fd).max() < 1e-7 * np.abs(fd).max()


def test_coaxial_table_m0_column(oracle):
    # (E|F)_{l0}^0 = sqrt(2l+1) f_l(kr) (G&D 4.83), f = j (R|R) or h (S|R)
    L, kr = 12, 3.3
    for kind in (0, 1):
        E = oracle.hel_coaxial_table(kind, L, kr)
        for l in range(L + 1):
            idx = l * (l + 1) * (l + 2) // 6
            f = spherical_jn(l, kr) + (1j * spherical_yn(l, kr) if kind else 0)
            assert abs(E[2 * idx] + 1j * E[2 * idx + 1] - math.sqrt(2 * l + 1) * f) < 1e-12 * max(1, abs(f))


def test_helmholtz_direct_golden(oracle):
    with open(os.path.join(GOLD, "helmholtz_direct.json")) as fh:
        g = json.load(fh)
    xs = np.array(g["xs"])
    q = np.array(g["q_re"]) + 1j * np.array(g["q_im"])
    d = np.array(g["d_re"]) + 1j * np.array(g["d_im"])
    f = oracle.hel_direct(g["k"], xs, q, np.array(g["xt"]), nrm=np.array(g["nrm"]), d=d)
    want = np.array(g["f_re"]) + 1j * np.array(g["f_im"])
    assert np.abs(f - want).max() < 1e-12 * np.abs(want).max()


@pytest.mark.parametrize("depth,p,tol", [(3, 8, 2e-4), (3, 12, 1e-5)])
def test_oracle_fmm_converges_to_direct(oracle, depth, p, tol):
    from conftest import fit_box
    rng = np.random.default_rng(3)
    n = 1500
    xs = rng.uniform(-1, 1, (n, 3))
    xt = rng.uniform(-1, 1, (400, 3))
    q = rng.normal(size=n) + 1j * rng.normal(size=n)
    nrm = rng.normal(size=(n, 3))
    nrm /= np.linalg.norm(nrm, axis=1)[:, None]
    d = rng.normal(size=n) + 1j * rng.normal(size=n)
    x0, D = fit_box(xs, xt)
    f, _, _ = oracle.fmm(2, xs, xt, x0, D, depth, [p] * 17, [p] * 17,
                         q=np.stack([q.real, q.imag], 1).copy(),
                         nrm=nrm, d=np.stack([d.real, d.imag], 1).copy(),
                         k=1.5, field=3, nthreads=4)
    fc = f[:, 0::2] + 1j * f[:, 1::2]
    fd = oracle.hel_direct(1.5, xs, q, xt, nrm=nrm, d=d, cutoff=1e-6)
    for c in range(4):
        e = np.linalg.norm(fc[:, c] - fd[:, c]) / np.linalg.norm(fd[:, c])
        assert e < tol, (c, e)
