"""Pins the CPU oracle (oracle/wbfmm_oracle.c) for the Laplace path.

The reference ships no golden vectors and cannot be built here (it needs
GLib and the absent blaswrap library), so every oracle function is pinned
against what the mathematics fixes: scipy's associated Legendre functions,
closed-form kernels summed directly (and in 50-digit mpmath for the golden
fixture), the identity of a rotation followed by its inverse, finite
differences for gradients, and the FMM's own convergence in the order p.
Each check is chosen so that a dropped term, a wrong sign, index or
transposed operand fails it.
"""

import json
import math
import os

import numpy as np
import pytest
from scipy.special import factorial, lpmv

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def test_legendre_matches_scipy(oracle):
    # Pbar_n^m = sqrt((2n+1)/(4pi) (n-m)!/(n+m)!) P_n^m without the
    # Condon-Shortley phase (src/util.c:67-116)
    for th in (0.0, 0.3, 1.2, 2.9, math.pi):
        N = 14
        P = oracle.legendre(N, th)
        for n in range(N + 1):
            for m in range(n + 1):
                ref = (np.sqrt((2 * n + 1) / (4 * np.pi) * factorial(n - m) / factorial(n + m))
                       * lpmv(m, n, np.cos(th)) * (-1) ** m)
                assert abs(P[n * (n + 1) // 2 + m] - ref) < 1e-13 * max(1.0, abs(ref))


def test_morton_and_lists_bruteforce(oracle):
    import ctypes
    L = oracle.lib()
    rng = np.random.default_rng(0)
    # Morton: x bit 0, y bit 1, z bit 2 of each triple (src/indexing.c:68-163)
    for _ in range(50):
        i, j, k = (int(v) for v in rng.integers(0, 1 << 20, 3))
        c = L.orc_morton_encode(i, j, k)
        ref = 0
        for b in range(20):
            ref |= ((i >> b) & 1) << (3 * b) | ((j >> b) & 1) << (3 * b + 1) | ((k >> b) & 1) << (3 * b + 2)
        assert c == ref
    # list 4 at level 3 by brute force over all boxes
    level = 3
    nb = 1 << level
    lst = (ctypes.c_uint64 * 216)()
    off = (ctypes.c_int * (3 * 216))()
    for b in range(0, 8 ** level, 37):
        n = L.orc_box_list4(level, b, lst, off)
        x, y, z = [sum(((b >> (3 * t + a)) & 1) << t for t in range(level)) for a in range(3)]
        want = set()
        for bb in range(8 ** level):
            xx, yy, zz = [sum(((bb >> (3 * t + a)) & 1) << t for t in range(level)) for a in range(3)]
            near = abs(xx - x) <= 1 and abs(yy - y) <= 1 and abs(zz - z) <= 1
            pnb = abs(xx // 2 - x // 2) <= 1 and abs(yy // 2 - y // 2) <= 1 and abs(zz // 2 - z // 2) <= 1
            if pnb and not near:
                want.add(bb)
        assert set(lst[:n]) == want
        assert n <= 189


def test_p2m_expansion_matches_direct(oracle):
    # expansion evaluated far away equals the direct sum
    # (tools/wbfmm-laplace-tests.c.in expansion test)
    rng = np.random.default_rng(1)
    N = 16
    src = rng.uniform(-0.3, 0.3, (20, 3))
    q = rng.normal(size=20)
    M = np.zeros((N + 1) ** 2)
    for s, qq in zip(src, q):
        oracle.lap_p2m(N, np.zeros(3), s, qq, M)
    for xf in ([1.3, -0.9, 1.1], [0.0, 0.0, 2.0], [-1.5, 0.2, -0.4]):
        xf = np.array(xf)
        fd = oracle.lap_direct(src, q, xf[None])[0, 0]
        fm = oracle.lap_m_eval(N, np.zeros(3), M, xf)
        assert abs(fm - fd) < 1e-9 * abs(fd)


def test_p2m_dipole_matches_direct(oracle):
    rng = np.random.default_rng(2)
    N = 16
    src = rng.uniform(-0.3, 0.3, (15, 3))
    nrm = rng.normal(size=(15, 3))
    nrm /= np.linalg.norm(nrm, axis=1)[:, None]
    d = rng.normal(size=15)
    M = np.zeros((N + 1) ** 2)
    for s, nn, dd in zip(src, nrm, d):
        oracle.lap_p2m_normal(N, np.zeros(3), s, nn, dd, M)
    xf = np.array([1.1, 1.2, -0.8])
    fd = oracle.lap_direct(src, None, xf[None], nrm=nrm, d=d)[0, 0]
    assert abs(oracle.lap_m_eval(N, np.zeros(3), M, xf) - fd) < 1e-9 * abs(fd)


@pytest.mark.parametrize("off", [(2, 0, 0), (-3, 1, 2), (0, 0, 2), (0, 0, -3),
                                 (2, -2, 3), (-2, -3, 0), (3, 3, 3), (-3, -3, -3)])
def test_m2l_rotate_coaxial_rotate(oracle, off):
    # rotate -> SR -> rotate back (src/pass-laplace.c:35-105) reproduces the
    # direct field at a target-box point, for axis, diagonal and generic
    # offsets (the three kinds of table entry in src/shifts.c)
    rng = np.random.default_rng(abs(hash(off)) % 1000)
    N = 18
    xs0 = np.zeros(3)
    xt0 = xs0 - np.array(off, float)
    src = xs0 + rng.uniform(-0.5, 0.5, (10, 3))
    q = rng.normal(size=10)
    M = np.zeros((N + 1) ** 2)
    for s, qq in zip(src, q):
        oracle.lap_p2m(N, xs0, s, qq, M)
    th, ph, ch, r = oracle.shift_angles(*off)
    H = oracle.H_rotation(N, th)
    Lc = oracle.lap_rotate(oracle.lap_coaxial(2, oracle.lap_rotate(M, N, H, ph, ch), N, N, r),
                           N, H, ch, ph)
    for _ in range(3):
        xf = xt0 + rng.uniform(-0.5, 0.5, 3)
        fl = oracle.lap_l_eval(N, xt0, Lc, xf)
        fd = oracle.lap_direct(src, q, xf[None])[0]
        assert np.abs(fl - fd).max() < 2e-5 * np.abs(fd).max()


def test_rotation_inverse_is_identity(oracle):
    # swapping phi and chi inverts the rotation (H(theta)^2 = I), which the
    # reference relies on for every rotate-back
    rng = np.random.default_rng(3)
    N = 20
    for th in (0.2, 1.1, 2.5, math.pi / 2, 2.9):
        H = oracle.H_rotation(N, th)
        Ci = rng.normal(size=(N + 1) ** 2)
        a = oracle.lap_rotate(Ci, N, H, -np.pi / 2, 0.6)
        b = oracle.lap_rotate(a, N, H, 0.6, -np.pi / 2)
        # H(theta) is an involution; the quad-precision recursion keeps
        # it to rounding even at N=20 (FP64 would give ~4e-5 here)
        assert np.abs(b - Ci).max() < 1e-13
        assert np.abs(a - Ci).max() > 1e-3


def test_rotation_preserves_field(oracle):
    # a rotated multipole evaluated at the rotated point equals the
    # original at the original point
    rng = np.random.default_rng(4)
    N = 14
    src = rng.uniform(-0.4, 0.4, (8, 3))
    q = rng.normal(size=8)
    M = np.zeros((N + 1) ** 2)
    for s, qq in zip(src, q):
        oracle.lap_p2m(N, np.zeros(3), s, qq, M)
    th, ph, ch, _ = oracle.shift_angles(-2, 1, -3)
    H = oracle.H_rotation(N, th)
    Mr = oracle.lap_rotate(M, N, H, ph, ch)
    # the new frame: z' along e = (2,-1,3), x' = (e_y,-e_x,0)/|.|
    e = np.array([2.0, -1.0, 3.0])
    ez = e / np.linalg.norm(e)
    ex = np.array([ez[1], -ez[0], 0.0])
    ex /= np.linalg.norm(ex)
    ey = np.cross(ez, ex)
    xf = np.array([1.7, -1.1, 0.9])
    xr = np.array([xf @ ex, xf @ ey, xf @ ez])
    f0 = oracle.lap_m_eval(N, np.zeros(3), M, xf)
    f1 = oracle.lap_m_eval(N, np.zeros(3), Mr, xr)
    assert abs(f0 - f1) < 1e-12 * abs(f0)


def test_local_gradient_matches_finite_difference(oracle):
    rng = np.random.default_rng(5)
    N = 10
    Lc = rng.normal(size=(N + 1) ** 2)
    xf = np.array([0.2, -0.3, 0.25])
    g = oracle.lap_l_eval(N, np.zeros(3), Lc, xf)[1:]
    h = 1e-6
    fd = [(oracle.lap_l_eval(N, np.zeros(3), Lc, xf + h * e, 1)[0]
           - oracle.lap_l_eval(N, np.zeros(3), Lc, xf - h * e, 1)[0]) / (2 * h) for e in np.eye(3)]
    assert np.allclose(g, fd, rtol=1e-7, atol=1e-9)


def test_coaxial_closed_forms(oracle):
    # SS/RR/SR coefficients against the factorial forms
    # (src/laplace.c:542-725); SR carries (-1)^(n+m)
    L = oracle.lib()
    f = math.factorial
    for n in range(0, 7):
        for nu in range(0, 7):
            for m in range(0, min(n, nu) + 1):
                if n >= nu:
                    ss = math.sqrt(f(n - m) * f(n + m) / (f(nu - m) * f(nu + m)) * (2 * nu + 1) / (2 * n + 1)) / f(n - nu)
                    assert abs(L.orc_lap_coaxial_cfft(0, n, nu, m) - ss) < 1e-12 * ss
                if nu >= n:
                    rr = math.sqrt(f(nu - m) * f(nu + m) / (f(n - m) * f(n + m)) * (2 * nu + 1) / (2 * n + 1)) / f(nu - n)
                    assert abs(L.orc_lap_coaxial_cfft(1, n, nu, m) - rr) < 1e-12 * rr
                sr = (-1) ** (n + m) * math.sqrt(f(n + nu) ** 2 / (f(n - m) * f(n + m) * f(nu - m) * f(nu + m))
                                                * (2 * nu + 1) / (2 * n + 1))
                assert abs(L.orc_lap_coaxial_cfft(2, n, nu, m) - sr) < 1e-12 * abs(sr)


def test_direct_sum_golden(oracle):
    # 50-digit mpmath sums committed under tests/golden
    with open(os.path.join(GOLD, "laplace_direct.json")) as fh:
        g = json.load(fh)
    xs = np.array(g["xs"])
    q = np.array(g["q"])
    nrm = np.array(g["nrm"])
    d = np.array(g["d"])
    xt = np.array(g["xt"])
    f = oracle.lap_direct(xs, q, xt, nrm=nrm, d=d)
    want = np.array(g["f"])
    assert np.abs(f - want).max() < 1e-13 * np.abs(want).max()


@pytest.mark.parametrize("depth,p,tol", [(3, 6, 2e-3), (3, 10, 1e-4), (4, 12, 2e-5)])
def test_oracle_fmm_converges_to_direct(oracle, depth, p, tol):
    from conftest import fit_box
    rng = np.random.default_rng(7)
    n = 3000
    xs = rng.uniform(0, 1, (n, 3))
    q = rng.normal(size=n)
    x0, D = fit_box(xs)
    f, _, _ = oracle.fmm(1, xs, xs, x0, D, depth, [p] * 17, [p] * 17, q=q, field=3)
    fd = oracle.lap_direct(xs, q, xs, cutoff=1e-6)
    for c in range(4):
        err = np.linalg.norm(f[:, c] - fd[:, c]) / np.linalg.norm(fd[:, c])
        assert err < tol, (c, err)


def test_oracle_fmm_dipoles_and_targets(oracle):
    from conftest import fit_box
    rng = np.random.default_rng(8)
    n = 1500
    xs = rng.uniform(-1, 1, (n, 3))
    xt = rng.uniform(-1, 1, (700, 3))
    q = rng.normal(size=n)
    nrm = rng.normal(size=(n, 3))
    nrm /= np.linalg.norm(nrm, axis=1)[:, None]
    d = rng.normal(size=n)
    x0, D = fit_box(xs, xt)
    f, _, _ = oracle.fmm(1, xs, xt, x0, D, 3, [12] * 17, [12] * 17, q=q, nrm=nrm, d=d, field=3)
    fd = oracle.lap_direct(xs, q, xt, nrm=nrm, d=d, cutoff=1e-6)
    for c in range(4):
        assert np.linalg.norm(f[:, c] - fd[:, c]) / np.linalg.norm(fd[:, c]) < 1e-4


# ---- second derivatives and the curl (vector sources) -------------------

@pytest.mark.parametrize("N", [2, 3, 6, 12, 20])
def test_local_second_derivatives_match_finite_difference(oracle, N):
    """orc_lap_l_hess against central differences of the local gradient
    (a wrong ladder sign or index fails at the 1e-3 level)"""
    rng = np.random.default_rng(100 + N)
    L = rng.normal(size=(N + 1) ** 2)
    x0 = np.array([0.2, -0.1, 0.3])
    for _ in range(3):
        xf = x0 + rng.uniform(-0.3, 0.3, 3)
        h = oracle.lap_l_hess(N, x0, L, xf)
        eps = 1e-5
        H = np.zeros((3, 3))
        for j in range(3):
            e = np.zeros(3)
            e[j] = eps
            H[:, j] = (oracle.lap_l_eval(N, x0, L, xf + e, 2)[1:4] -
                       oracle.lap_l_eval(N, x0, L, xf - e, 2)[1:4]) / (2 * eps)
        fd = np.array([H[0, 0], H[1, 1], H[2, 2], H[0, 1], H[1, 2], H[2, 0]])
        assert np.abs(h - fd).max() < 1e-7 * np.abs(fd).max()
        # the field is harmonic: the trace vanishes
        assert abs(h[0] + h[1] + h[2]) < 1e-12 * np.abs(h).max()


def test_direct_second_derivatives_match_finite_difference(oracle):
    rng = np.random.default_rng(11)
    xs = rng.uniform(0, 1, (50, 3))
    q = rng.normal(size=50)
    xt = rng.uniform(1.5, 2.0, (6, 3))
    f = oracle.lap_direct(xs, q, xt, field=7)
    eps = 1e-5
    for c, (a, b) in enumerate([(0, 0), (1, 1), (2, 2), (0, 1), (1, 2), (2, 0)]):
        e = np.zeros(3)
        e[b] = eps
        fd = (oracle.lap_direct(xs, q, xt + e)[:, 1 + a] -
              oracle.lap_direct(xs, q, xt - e)[:, 1 + a]) / (2 * eps)
        assert np.abs(f[:, 4 + c] - fd).max() < 1e-6 * np.abs(f[:, 4:]).max()


def reference_curl(xs, w, xt, grad):
    """the reference's direct curl (src/direct-laplace.c:87-198), written
    out: grad G x w and d_j grad G x w with G = 1/(4 pi r)"""
    out = np.zeros((len(xt), 12 if grad else 3))
    for i, x in enumerate(xt):
        for y, s in zip(xs, w):
            dr = x - y
            r = np.linalg.norm(dr)
            r3 = r ** 3 * 4 * np.pi
            r5 = r3 * r * r
            nR = -dr / r3
            out[i, :3] += np.cross(nR, s)
            if grad:
                for j in range(3):
                    row = 3 * dr[j] * dr / r5
                    row[j] -= 1.0 / r3
                    out[i, 3 + 3 * j:6 + 3 * j] += np.cross(row, s)
    return out


def test_curl_from_components_matches_reference_curl(oracle):
    """the combination of the three components' gradients and second
    derivatives equals the reference's closed-form curl and curl-gradient"""
    rng = np.random.default_rng(12)
    xs = rng.uniform(0, 1, (40, 3))
    w = rng.normal(size=(40, 3))
    xt = rng.uniform(-0.5, 1.5, (9, 3))
    rows = [oracle.lap_direct(xs, w[:, c], xt, field=7) for c in range(3)]
    got = oracle.curl_from_components(rows)
    want = reference_curl(xs, w, xt, True)
    assert np.abs(got - want).max() < 1e-12 * np.abs(want).max()


@pytest.mark.parametrize("depth,p", [(3, 10), (4, 12)])
def test_oracle_fmm_curl_converges_to_direct(oracle, depth, p):
    from conftest import fit_box
    rng = np.random.default_rng(13 + depth)
    n = 2500
    xs = rng.uniform(0, 1, (n, 3))
    w = rng.normal(size=(n, 3))
    x0, D = fit_box(xs)
    rows = [oracle.fmm(1, xs, xs, x0, D, depth, [p] * 17, [p] * 17, q=w[:, c], field=7)[0]
            for c in range(3)]
    got = oracle.curl_from_components(rows)
    drows = [oracle.lap_direct(xs, w[:, c], xs, field=7, cutoff=1e-6) for c in range(3)]
    want = oracle.curl_from_components(drows)
    for c in range(12):
        err = np.linalg.norm(got[:, c] - want[:, c]) / np.linalg.norm(want[:, c])
        assert err < 1e-4, (c, err)
