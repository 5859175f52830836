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
- the one-wavefront-per-box FP64 field kernel against the thread kernel;
- the class-packed FP32 M2L (forced at every level) against the chunked
  kernel, the FP64 solve and the direct sum.
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


@pytest.mark.parametrize("p,k", [(10, 3.0), (23, 1.5)])
def test_packed_fp32_m2l_matches_chunked(wb, monkeypatch, p, k):
    # the class-packed FP32 M2L (forced at every level) against the chunked
    # kernel and the FP64 solve on the same tree: FP32 rounding, summed in
    # another order
    xs, q, xt = cube(730 + p, 4000, 1000)
    x0, D = fit_box(xs, xt)
    monkeypatch.setenv("WBFMM_HEL_FP64_KA", "100")
    monkeypatch.setenv("WBFMM_HEL_PACKED", "0")
    fc = solve(wb, xs, q, None, xt, x0, D, 4, [p] * 17, k, 32)
    monkeypatch.setenv("WBFMM_HEL_PACKED", "1")
    fp = solve(wb, xs, q, None, xt, x0, D, 4, [p] * 17, k, 32)
    monkeypatch.delenv("WBFMM_HEL_PACKED")
    monkeypatch.delenv("WBFMM_HEL_FP64_KA")
    f64 = solve(wb, xs, q, None, xt, x0, D, 4, [p] * 17, k, 64)
    r, r64 = colrel(fp, fc), colrel(fp, f64)
    print(f"p={p}: packed vs chunked {r.max():.2e}, packed vs FP64 {r64.max():.2e}", flush=True)
    assert r.max() < TOL and r64.max() < TOL
    assert not np.array_equal(fp, fc)  # the packed kernel ran


def test_packed_fp32_m2l_sphere_with_dipoles(wb, oracle, monkeypatch):
    # the mixed sphere of the boundary test with the packed FP32 M2L forced
    # at its FP32 levels: against the FP64 solve and the direct sum
    n, k, depth = 20_000, 16.0, 5
    xs, q, d = sphere(721, n)
    x0, D = fit_box(xs)
    orders = trunc_orders(wb, D, k, 1e-6, depth)
    monkeypatch.setenv("WBFMM_HEL_PACKED", "1")
    fm = solve(wb, xs, q, d, xs, x0, D, depth, orders, k, 32)
    monkeypatch.delenv("WBFMM_HEL_PACKED")
    f64 = solve(wb, xs, q, d, xs, x0, D, depth, orders, k, 64)
    r = colrel(fm, f64)
    print(f"sphere packed mixed vs FP64: max rel diff per component {r.max():.2e}", flush=True)
    assert r.max() < TOL, r
    idx = np.arange(0, n, 97)
    fd = oracle.hel_direct(k, xs, q, xs[idx], field=1, nrm=xs, d=d, cutoff=1e-6)
    fc = fm[idx, 0] + 1j * fm[idx, 1]
    e = np.linalg.norm(fc - fd[:, 0]) / np.linalg.norm(fd[:, 0])
    print(f"sphere packed mixed: potential rel L2 error {e:.2e}", flush=True)
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


def _hel_f_api(wb):
    """ctypes signatures of the Helmholtz single-precision entry points"""
    L = wb.lib()
    vp, i, u, fl = ctypes.c_void_p, ctypes.c_int, ctypes.c_uint, ctypes.c_float
    sz = ctypes.c_size_t
    sigs = {
        "wbfmm_tree_new_f": (ctypes.POINTER(wb.WbfmmTree), [vp, fl, u]),
        "wbfmm_tree_add_points_f": (i, [vp, vp, sz, vp, sz, u, i]),
        "wbfmm_tree_refine_f": (i, [vp]),
        "wbfmm_tree_coefficient_init_f": (i, [vp, u, u, u]),
        "wbfmm_tree_leaf_expansions_f": (i, [vp, fl, vp, i, vp, i, i, vp]),
        "wbfmm_shift_angle_table_init_f": (i, []),
        "wbfmm_shift_operators_new_f": (vp, [u, i, vp]),
        "wbfmm_shift_operators_coaxial_SR_init_f": (i, [vp, fl, u, u, fl, vp]),
        "wbfmm_shift_operators_coaxial_SS_init_f": (i, [vp, fl, u, u, fl, vp]),
        "wbfmm_upward_pass_f": (i, [vp, vp, u, vp]),
        "wbfmm_downward_pass_ref_f": (i, [vp, vp, u, vp, i]),
        "wbfmm_tree_box_local_field_f": (i, [vp, u, u, fl, vp, vp, i, vp, i, vp, i, i, u, vp]),
        "wbfmm_target_list_new_f": (vp, [vp, u]),
        "wbfmm_target_list_add_points_f": (i, [vp, vp, sz, u]),
        "wbfmm_target_list_coefficients_init": (i, [vp, u]),
        "wbfmm_target_list_local_coefficients_f": (i, [vp, fl, vp]),
        "wbfmm_target_list_local_field_f": (i, [vp, vp, i, vp, i, vp, i]),
        "wbfmm_target_list_free": (i, [vp]),
        "wbfmm_total_field_f": (i, [fl, vp, i, vp, i, vp, i, vp, i, i, i, vp, vp, i]),
        "wbfmm_total_field_grad_f": (i, [fl, vp, i, vp, i, vp, i, vp, i, i, i, vp, vp, i]),
    }
    for name, (res, args) in sigs.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    return L


def test_single_precision_helmholtz_api(wb, oracle):
    """the float Helmholtz entry points (src/wbfmm.h:938-951, :997-1003,
    :1059-1082, :1102-1104) run the mixed-precision tree: with the same
    float inputs the target-list field equals the double entry points'
    on a mixed tree, rounded to float; the box field and the direct sums
    agree with the double ones to float rounding"""
    L = _hel_f_api(wb)
    Pf = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    rng = np.random.default_rng(777)
    n, nt, depth, p, k = 3000, 400, 4, 12, 3.0
    v = rng.normal(size=(n, 3))
    xs = (v / np.linalg.norm(v, axis=1)[:, None]).astype(np.float32)
    xt = (0.8 * xs[::7]).astype(np.float32)[:nt]
    nt = len(xt)
    q = (rng.normal(size=(n, 2))).astype(np.float32)
    d = (rng.normal(size=(n, 2))).astype(np.float32)
    x0 = np.full(3, -1.02, np.float32)
    D = np.float32(2.04)
    kf = np.float32(k)

    tp = L.wbfmm_tree_new_f(Pf(x0), D, n)
    t = ctypes.cast(tp, ctypes.c_void_p).value
    assert L.wbfmm_tree_add_points_f(t, Pf(xs), 12, Pf(xs), 12, n, 0) == 0
    for _ in range(depth):
        L.wbfmm_tree_refine_f(t)
    tp.contents.problem = 2
    tp.contents.nq = 1
    for lv in range(1, depth + 1):
        L.wbfmm_tree_coefficient_init_f(t, lv, p, p)
    L.wbfmm_shift_angle_table_init_f()
    op = L.wbfmm_shift_operators_new_f(p, 0, None)
    for lv in range(2, depth + 1):
        assert L.wbfmm_shift_operators_coaxial_SR_init_f(op, D, lv, p, kf, None) == 0
    for lv in range(3, depth + 1):
        assert L.wbfmm_shift_operators_coaxial_SS_init_f(op, D, lv, p, kf, None) == 0
    tl = L.wbfmm_target_list_new_f(t, nt)
    L.wbfmm_target_list_add_points_f(tl, Pf(xt), 12, nt)
    L.wbfmm_tree_leaf_expansions_f(t, kf, Pf(q), 2, Pf(d), 2, 1, None)
    for lv in range(depth, 2, -1):
        L.wbfmm_upward_pass_f(t, op, lv, None)
    for lv in range(2, depth + 1):
        L.wbfmm_downward_pass_ref_f(t, op, lv, None, 0)
    L.wbfmm_target_list_coefficients_init(tl, 3)
    L.wbfmm_target_list_local_coefficients_f(tl, kf, None)
    ff = np.zeros((nt, 8), np.float32)
    L.wbfmm_target_list_local_field_f(tl, Pf(q), 2, Pf(d), 2, Pf(ff), 8)
    x5 = np.ascontiguousarray(xt[5])
    fb = np.zeros(2, np.float32)
    L.wbfmm_tree_box_local_field_f(t, depth, 0, kf, Pf(x5), Pf(fb), 2, Pf(q), 2, Pf(d), 2, 1, 1,
                                   None)
    fdp = np.zeros(2, np.float32)
    fdg = np.zeros(6, np.float32)
    L.wbfmm_total_field_f(kf, Pf(xs), 3, Pf(q), 2, Pf(xs), 3, Pf(d), 2, 1, n, Pf(x5), Pf(fdp), 1)
    L.wbfmm_total_field_grad_f(kf, Pf(xs), 3, Pf(q), 2, Pf(xs), 3, Pf(d), 2, 1, n, Pf(x5),
                               Pf(fdg), 1)
    wb.sync()
    L.wbfmm_target_list_free(tl)
    wb.lib().wbfmm_shift_operators_free(op)
    wb.lib().wbfmm_tree_free(t)

    # the double entry points on a mixed tree with the same (float) inputs
    x64, q64, d64 = xs.astype(np.float64), q.astype(np.float64), d.astype(np.float64)
    xt64 = xt.astype(np.float64)
    S = wb.HelmholtzSolver(np.ascontiguousarray(x64), x0.astype(np.float64), float(D), depth,
                           [p] * 17, [p] * 17, float(kf), normals=np.ascontiguousarray(x64),
                           targets=np.ascontiguousarray(xt64), precision=32)
    S.build_tree()
    S.passes(np.ascontiguousarray(q64), d=np.ascontiguousarray(d64))
    Ld = wb.lib()
    tl2 = Ld.wbfmm_target_list_new(S.t, nt)
    Ld.wbfmm_target_list_add_points(tl2, wb.ptr(np.ascontiguousarray(xt64)), 24, nt)
    Ld.wbfmm_target_list_coefficients_init(tl2, 3)
    Ld.wbfmm_target_list_local_coefficients.argtypes = [ctypes.c_void_p, ctypes.c_double,
                                                        ctypes.c_void_p]
    Ld.wbfmm_target_list_local_coefficients(tl2, float(kf), None)
    g = np.zeros((nt, 8))
    Ld.wbfmm_target_list_local_field(tl2, wb.ptr(np.ascontiguousarray(q64)), 2,
                                     wb.ptr(np.ascontiguousarray(d64)), 2, wb.ptr(g), 8)
    wb.sync()
    Ld.wbfmm_target_list_free(tl2)
    S.close()
    assert np.array_equal(ff, g.astype(np.float32))
    assert np.abs(fb - g[5, :2]).max() <= 2e-6 * np.abs(g[5, :2]).max()
    zq, zd = q64[:, 0] + 1j * q64[:, 1], d64[:, 0] + 1j * d64[:, 1]
    fd = oracle.hel_direct(float(kf), x64, zq, x5[None].astype(np.float64), field=3, nrm=x64, d=zd)
    assert np.allclose(fdp, [fd[0, 0].real, fd[0, 0].imag], rtol=2e-6, atol=0)
    want_g = np.array([fd[0, 1].real, fd[0, 1].imag, fd[0, 2].real, fd[0, 2].imag,
                       fd[0, 3].real, fd[0, 3].imag])
    assert np.abs(fdg - want_g).max() <= 2e-6 * np.abs(want_g).max()
