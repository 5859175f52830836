"""Mixed-precision Helmholtz (wbfmm_tree_set_precision(t, 32) on a
Helmholtz tree) on the GPU.

Levels with k a_l <= 2 hold FP32 coefficients scaled per degree (M_n / s_n,
L_n s_n, s_n = min(1, (k a_l)^n / (2n+1)!!)) and run FP32 matrix-core
passes and an FP32 near field with compensated sums; the high-frequency
levels stay FP64.  Checked here:

- with every level forced to FP64 (WBFMM_HEL_FP64_KA large) the mixed
  dispatch gives the FP64 path's bits;
- all-FP32 and mixed FP64/FP32 solves against the FP64 solve on the same
  tree, relative to the largest value of each component (FP32 rounding,
  6e-8 per stored value, grows through a few hundred terms per
  translation: 2e-6 is the bound);
- their error against the FP64 direct sum stays within 1e-6 where the FP64
  path's truncation error is well below it;
- the coefficients of FP32 levels, downloaded in the reference's scaling,
  against the FP64 tree's;
- the one-wavefront-per-box FP64 field kernel against the thread kernel.
"""

import ctypes

import numpy as np
import pytest

from conftest import fit_box

pytestmark = pytest.mark.gpu

TOL = 2e-6


def c2(z):
    return np.ascontiguousarray(np.stack([np.real(z), np.imag(z)], 1))


def solve(W, xs, q, d, xt, x0, D, depth, orders, k, precision, keep=False):
    S = W.HelmholtzSolver(np.ascontiguousarray(xs), x0, D, depth, orders, orders, k,
                          normals=None if d is None else np.ascontiguousarray(xs),
                          targets=np.ascontiguousarray(xt), precision=precision)
    S.build_tree()
    S.passes(None if q is None else c2(q), d=None if d is None else c2(d))
    f = np.zeros((len(xt), 8))
    S.field(np.ascontiguousarray(xt), None if q is None else c2(q), f, field=3, fstr=8,
            d=None if d is None else c2(d))
    W.sync()
    if keep:
        return f, S
    S.close()
    return f


def colrel(a, b):
    return np.array([np.abs(a[:, c] - b[:, c]).max() / np.abs(b[:, c]).max()
                     for c in range(a.shape[1])])


def trunc_orders(W, D, k, tol, depth):
    t = W.WbfmmTree()
    t.D = D
    o = [W.lib().wbfmm_truncation_number(ctypes.byref(t), k, max(l, 2), tol)
         for l in range(depth + 1)]
    return o + [o[-1]] * (17 - len(o))


def cube(seed, n, nt=None):
    rng = np.random.default_rng(seed)
    xs = rng.uniform(-1, 1, (n, 3))
    q = rng.normal(size=n) + 1j * rng.normal(size=n)
    xt = xs if nt is None else rng.uniform(-1, 1, (nt, 3))
    return xs, q, xt


def sphere(seed, n):
    rng = np.random.default_rng(seed)
    v = rng.normal(size=(n, 3))
    xs = v / np.linalg.norm(v, axis=1)[:, None]
    half = np.arange(n) % 2 == 0
    q = np.where(half, rng.normal(size=n) + 1j * rng.normal(size=n), 0.0)
    d = np.where(~half, rng.normal(size=n) + 1j * rng.normal(size=n), 0.0)
    return xs, q, d


def test_every_level_fp64_gives_fp64_bits(wb, monkeypatch):
    xs, q, xt = cube(701, 3000, 800)
    x0, D = fit_box(xs, xt)
    # the threshold below every level's k a_l: every level runs FP64
    monkeypatch.setenv("WBFMM_HEL_FP64_KA", "-1")
    f_mix = solve(wb, xs, q, None, xt, x0, D, 3, [10] * 17, 3.0, 32)
    monkeypatch.delenv("WBFMM_HEL_FP64_KA")
    f64 = solve(wb, xs, q, None, xt, x0, D, 3, [10] * 17, 3.0, 64)
    assert np.array_equal(f_mix, f64)


@pytest.mark.parametrize("p,k", [(6, 1.0), (10, 3.0), (16, 2.0), (23, 1.5)])
def test_all_fp32_levels_match_fp64(wb, monkeypatch, p, k):
    # the threshold above every level's k a_l: every level runs FP32
    xs, q, xt = cube(710 + p, 4000, 1000)
    x0, D = fit_box(xs, xt)
    monkeypatch.setenv("WBFMM_HEL_FP64_KA", "100")
    f32 = solve(wb, xs, q, None, xt, x0, D, 4, [p] * 17, k, 32)
    monkeypatch.delenv("WBFMM_HEL_FP64_KA")
    f64 = solve(wb, xs, q, None, xt, x0, D, 4, [p] * 17, k, 64)
    r = colrel(f32, f64)
    print(f"p={p} k={k}: max rel diff per component {r.max():.2e}", flush=True)
    assert r.max() < TOL, r
    assert not np.array_equal(f32, f64)  # the FP32 path ran


def test_mixed_sphere_boundary_levels(wb, oracle):
    # k = 16 on the unit sphere, orders from the truncation number at 1e-6:
    # levels 2 and 3 (k a_l = 7, 3.5) stay FP64, levels 4 and 5 run FP32;
    # the upward and downward passes cross between the two
    n, k, depth = 20_000, 16.0, 5
    xs, q, d = sphere(720, n)
    x0, D = fit_box(xs)
    orders = trunc_orders(wb, D, k, 1e-6, depth)
    fm = solve(wb, xs, q, d, xs, x0, D, depth, orders, k, 32)
    f64 = solve(wb, xs, q, d, xs, x0, D, depth, orders, k, 64)
    r = colrel(fm, f64)
    print(f"sphere mixed vs FP64: max rel diff per component {r.max():.2e}", flush=True)
    assert r.max() < TOL, r
    # error against the FP64 direct sum on sampled points, potential
    idx = np.arange(0, n, 97)
    fd = oracle.hel_direct(k, xs, q, xs[idx], field=1, nrm=xs, d=d, cutoff=1e-6)
    for name, f in (("mixed", fm), ("fp64", f64)):
        fc = f[idx, 0] + 1j * f[idx, 1]
        e = np.linalg.norm(fc - fd[:, 0]) / np.linalg.norm(fd[:, 0])
        print(f"sphere {name}: potential rel L2 error {e:.2e}", flush=True)
        assert e < 1e-6


def test_mixed_target_list_and_dipoles(wb):
    # the target-list field of a mixed tree equals the batched field
    xs, q, d = sphere(730, 6000)
    xt = np.ascontiguousarray(0.9 * xs[::3])
    x0, D = fit_box(xs)
    k = 6.0
    orders = trunc_orders(wb, D, k, 1e-5, 4)
    f, S = solve(wb, xs, q, d, xt, x0, D, 4, orders, k, 32, keep=True)
    L = wb.lib()
    tl = L.wbfmm_target_list_new(S.t, len(xt))
    L.wbfmm_target_list_add_points(tl, wb.ptr(xt), 24, len(xt))
    L.wbfmm_target_list_coefficients_init(tl, 3)
    L.wbfmm_target_list_local_coefficients.argtypes = [ctypes.c_void_p, ctypes.c_double,
                                                       ctypes.c_void_p]
    L.wbfmm_target_list_local_coefficients(tl, k, None)
    g = np.zeros((len(xt), 8))
    qq, dd = c2(q), c2(d)
    L.wbfmm_target_list_local_field(tl, wb.ptr(qq), 2, wb.ptr(dd), 2, wb.ptr(g), 8)
    wb.sync()
    L.wbfmm_target_list_free(tl)
    S.close()
    assert np.abs(g - f).max() <= 1e-12 * np.abs(f).max()


def test_mixed_coefficients_download(wb, monkeypatch):
    xs, q, xt = cube(740, 3000)
    x0, D = fit_box(xs)
    k, p, depth = 2.0, 12, 4
    monkeypatch.setenv("WBFMM_HEL_FP64_KA", "100")
    _, S32 = solve(wb, xs, q, None, xt, x0, D, depth, [p] * 17, k, 32, keep=True)
    monkeypatch.delenv("WBFMM_HEL_FP64_KA")
    _, S64 = solve(wb, xs, q, None, xt, x0, D, depth, [p] * 17, k, 64, keep=True)
    try:
        for lvl in (2, 3, 4):
            for which in (0, 1):
                a = S32.coefficients(lvl, which)
                b = S64.coefficients(lvl, which)
                C = (p + 1) ** 2
                n = np.floor(np.sqrt(np.arange(C))).astype(int)
                for deg in range(p + 1):
                    sel = n == deg
                    den = np.abs(b[:, sel]).max()
                    if den > 0:
                        assert np.abs(a[:, sel] - b[:, sel]).max() <= 1e-5 * den, (lvl, which, deg)
    finally:
        S32.close()
        S64.close()


def test_box_field_kernel_fp64_matches_thread_kernel(wb, monkeypatch):
    xs, q, d = sphere(750, 8000)
    xt = np.ascontiguousarray(np.concatenate([xs[::2], 0.5 * xs[::7]]))
    x0, D = fit_box(xs)
    k = 5.0
    orders = trunc_orders(wb, D, k, 1e-6, 4)
    f_thr = solve(wb, xs, q, d, xt, x0, D, 4, orders, k, 64)
    monkeypatch.setenv("WBFMM_HEL_FIELD", "box")
    f_box = solve(wb, xs, q, d, xt, x0, D, 4, orders, k, 64)
    r = colrel(f_box, f_thr)
    assert r.max() < 1e-12, r
