"""FP32 Laplace path (wbfmm_tree_set_precision(t, 32)) on the GPU.

The FP32 matrix-core operand layout against numpy; FP32 solves against
FP64 solves on the same tree (orders 2..15, dipoles, a deep tree where
unscaled coefficients would leave the FP32 range), the FP32 coefficients
against the FP64 ones through the download, and the FP32 error against
the FP64 direct sum.  Tolerances are relative to the largest value of
each field component: FP32 coefficients carry 6e-8 relative rounding per
stored value and the translations add up a few hundred terms each."""

import numpy as np
import pytest

from conftest import fit_box

pytestmark = pytest.mark.gpu

TOL_POT = 1e-5
TOL_GRAD = 3e-5


def solve(W, xs, q, xt, x0, D, depth, p, precision, nrm=None, d=None, keep=False):
    S = W.LaplaceSolver(np.ascontiguousarray(xs), x0, D, depth, [p] * 17, [p] * 17,
                        targets=np.ascontiguousarray(xt), host_mirror=False,
                        normals=None if nrm is None else np.ascontiguousarray(nrm),
                        precision=precision)
    f = np.zeros((len(xt), 4))
    S.solve(np.ascontiguousarray(q), f, field=3, fstr=4,
            d=None if d is None else np.ascontiguousarray(d))
    W.sync()
    if keep:
        return f, S
    S.close()
    return f


def colrel(a, b):
    return [np.abs(a[:, c] - b[:, c]).max() / np.abs(b[:, c]).max() for c in range(a.shape[1])]


def test_mfma_f32_operand_layout(wb):
    # the FP32 kernels assume lane l supplies A[l%16][l/16], B[l/16][l%16]
    # and receives D[4(l/16) + q][l%16]
    out = np.zeros(256, dtype=np.float32)
    assert wb.lib().wbfmm_hip_mfma_layout_probe_f32(wb.ptr(out)) == 0
    A = 1.0 + np.arange(16)[:, None] + 100.0 * np.arange(4)[None, :]
    Bm = 1.0 + 3.0 * np.arange(4)[:, None] + 1000.0 * np.arange(16)[None, :]
    Dm = A @ Bm
    got = np.zeros((16, 16))
    for l in range(64):
        for q in range(4):
            got[4 * (l // 16) + q, l % 16] = out[4 * l + q]
    assert np.array_equal(got, Dm)


@pytest.mark.parametrize("p", [2, 5, 8, 12, 15])
def test_fp32_matches_fp64_same_tree(wb, p):
    rng = np.random.default_rng(300 + p)
    xs = rng.uniform(0, 1, (6000, 3))
    q = rng.normal(size=6000)
    xt = rng.uniform(0, 1, (1500, 3))
    x0, D = fit_box(xs, xt)
    f64 = solve(wb, xs, q, xt, x0, D, 4, p, 64)
    f32 = solve(wb, xs, q, xt, x0, D, 4, p, 32)
    e = colrel(f32, f64)
    assert e[0] < TOL_POT and max(e[1:]) < TOL_GRAD, e


def test_fp32_dipoles_and_deep_tree(wb):
    # depth 7 with clustered points: level-7 boxes of width ~1/128, where
    # the unscaled M2L factors reach w^-(2p+1) ~ 1e45 at p = 12
    rng = np.random.default_rng(310)
    n = 20000
    xs = 0.5 + 0.08 * rng.normal(size=(n, 3))
    nrm = rng.normal(size=(n, 3))
    nrm /= np.linalg.norm(nrm, axis=1)[:, None]
    q = rng.normal(size=n)
    d = rng.normal(size=n)
    x0, D = fit_box(xs)
    f64 = solve(wb, xs, q, xs, x0, D, 7, 12, 64, nrm=nrm, d=d)
    f32 = solve(wb, xs, q, xs, x0, D, 7, 12, 32, nrm=nrm, d=d)
    assert np.isfinite(f32).all()
    e = colrel(f32, f64)
    assert e[0] < TOL_POT and max(e[1:]) < TOL_GRAD, e


def test_fp32_coefficients_download(wb):
    rng = np.random.default_rng(320)
    xs = rng.uniform(-1, 1, (5000, 3))
    q = rng.normal(size=5000)
    x0, D = fit_box(xs)
    _, S64 = solve(wb, xs, q, xs, x0, D, 4, 10, 64, keep=True)
    _, S32 = solve(wb, xs, q, xs, x0, D, 4, 10, 32, keep=True)
    for level in (2, 3, 4):
        for which in (0, 1):
            a = S64.coefficients(level, which)
            b = S32.coefficients(level, which)
            # per degree: coefficients of different degrees differ by
            # powers of the box width
            for n in range(11):
                sl = slice(n * n, (n + 1) * (n + 1))
                ref = np.abs(a[:, sl]).max()
                assert np.abs(b[:, sl] - a[:, sl]).max() <= 1e-5 * ref, (level, which, n)
    S64.close()
    S32.close()


def test_fp32_error_against_direct(wb, oracle):
    # tol 1e-4 class order (8): the FP32 solve's error against the FP64
    # direct sum is the truncation error, not the rounding
    rng = np.random.default_rng(330)
    xs = rng.uniform(0, 1, (8000, 3))
    q = rng.normal(size=8000)
    x0, D = fit_box(xs)
    f64 = solve(wb, xs, q, xs, x0, D, 4, 8, 64)
    f32 = solve(wb, xs, q, xs, x0, D, 4, 8, 32)
    idx = rng.choice(8000, 300, replace=False)
    fd = oracle.lap_direct(xs, q, xs[idx], 3, cutoff=1e-6)
    e64 = np.linalg.norm(f64[idx, 0] - fd[:, 0]) / np.linalg.norm(fd[:, 0])
    e32 = np.linalg.norm(f32[idx, 0] - fd[:, 0]) / np.linalg.norm(fd[:, 0])
    assert e64 < 1e-4 and e32 < 1e-4
    assert abs(e32 - e64) < 0.1 * e64 + 1e-6, (e32, e64)


def _f_api(wb):
    """ctypes signatures of the single-precision entry points"""
    import ctypes
    L = wb.lib()
    vp, i, u, f = ctypes.c_void_p, ctypes.c_int, ctypes.c_uint, ctypes.c_float
    sz = ctypes.c_size_t
    sigs = {
        "wbfmm_tree_new_f": (ctypes.POINTER(wb.WbfmmTree), [vp, f, u]),
        "wbfmm_tree_add_points_f": (i, [vp, vp, sz, vp, sz, u, i]),
        "wbfmm_tree_refine_f": (i, [vp]),
        "wbfmm_tree_laplace_coefficient_init_f": (i, [vp, u, u, u]),
        "wbfmm_tree_laplace_leaf_expansions_f": (i, [vp, vp, i, vp, i, i, vp]),
        "wbfmm_shift_angle_table_init_f": (i, []),
        "wbfmm_shift_operators_new_f": (vp, [u, i, vp]),
        "wbfmm_laplace_coaxial_translate_init_f": (i, [i]),
        "wbfmm_laplace_upward_pass_f": (i, [vp, vp, u, vp]),
        "wbfmm_laplace_downward_pass_ref_f": (i, [vp, vp, u, vp, i]),
        "wbfmm_laplace_box_field_f": (i, [vp, u, u, vp, i, vp, i, u, i, vp, vp, i, vp]),
        "wbfmm_laplace_field_direct_f": (i, [vp, i, vp, i, i, vp, i, vp, i, i, u, vp, vp, i]),
        "wbfmm_target_list_new_f": (vp, [vp, u]),
        "wbfmm_target_list_add_points_f": (i, [vp, vp, sz, u]),
        "wbfmm_laplace_target_list_local_coefficients_f": (i, [vp, u, vp]),
        "wbfmm_target_list_local_field_f": (i, [vp, vp, i, vp, i, vp, i]),
        "wbfmm_target_list_coefficients_init": (i, [vp, u]),
        "wbfmm_target_list_free": (i, [vp]),
    }
    for name, (res, args) in sigs.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    return L


def test_single_precision_api(wb, oracle):
    # the float entry points (src/wbfmm.h:432-440, :690-740, :996-1112) run
    # the FP32 tree: with the same float inputs they give the bits of the
    # double entry points on an FP32 tree, rounded to float
    import ctypes
    L = _f_api(wb)
    P = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    rng = np.random.default_rng(91)
    n, nt, depth, p = 4000, 700, 4, 9
    xs = rng.uniform(0, 1, (n, 3)).astype(np.float32)
    xt = rng.uniform(0, 1, (nt, 3)).astype(np.float32)
    q = rng.normal(size=n).astype(np.float32)
    x0 = np.full(3, -0.01, np.float32)
    D = np.float32(1.02)

    tp = L.wbfmm_tree_new_f(P(x0), D, n)
    t = ctypes.cast(tp, ctypes.c_void_p).value
    assert L.wbfmm_tree_add_points_f(t, P(xs), 12, None, 0, n, 0) == 0
    for _ in range(depth):
        L.wbfmm_tree_refine_f(t)
    tp.contents.problem = 1
    tp.contents.nq = 1
    for lv in range(1, depth + 1):
        L.wbfmm_tree_laplace_coefficient_init_f(t, lv, p, p)
    L.wbfmm_shift_angle_table_init_f()
    op = L.wbfmm_shift_operators_new_f(p, 0, None)
    L.wbfmm_laplace_coaxial_translate_init_f(p + 2)
    L.wbfmm_tree_laplace_leaf_expansions_f(t, P(q), 1, None, 1, 1, None)
    for lv in range(depth, 2, -1):
        L.wbfmm_laplace_upward_pass_f(t, op, lv, None)
    for lv in range(2, depth + 1):
        L.wbfmm_laplace_downward_pass_ref_f(t, op, lv, None, 0)
    tl = L.wbfmm_target_list_new_f(t, nt)
    L.wbfmm_target_list_add_points_f(tl, P(xt), 12, nt)
    L.wbfmm_target_list_coefficients_init(tl, 3)
    L.wbfmm_laplace_target_list_local_coefficients_f(tl, 1, None)
    ff = np.zeros((nt, 4), np.float32)
    L.wbfmm_target_list_local_field_f(tl, P(q), 1, None, 1, P(ff), 4)
    # one target through the box field, potential only
    fb = np.zeros(1, np.float32)
    x5 = np.ascontiguousarray(xt[5])
    L.wbfmm_laplace_box_field_f(t, depth, 0, P(q), 1, None, 1, 1, 1, P(x5), P(fb), 1, None)
    # direct sum at the same target
    fdr = np.zeros(3, np.float32)
    L.wbfmm_laplace_field_direct_f(P(xs), 3, None, 3, n, P(q), 1, None, 1, 1, 2, P(x5),
                                   P(fdr), 3)
    wb.sync()
    L.wbfmm_target_list_free(tl)
    lib = wb.lib()
    lib.wbfmm_shift_operators_free(op)
    lib.wbfmm_tree_free(t)

    f64 = solve(wb, xs.astype(np.float64), q.astype(np.float64), xt.astype(np.float64),
                x0.astype(np.float64), float(D), depth, p, 32)
    assert np.array_equal(ff, f64.astype(np.float32))
    f1 = solve(wb, xs.astype(np.float64), q.astype(np.float64), xt[5:6].astype(np.float64),
               x0.astype(np.float64), float(D), depth, p, 32)
    assert abs(fb[0] - f1[0, 0]) <= 1e-6 * abs(f1[0, 0])
    fd = oracle.lap_direct(xs.astype(np.float64), q.astype(np.float64), x5[None].astype(np.float64),
                           2)
    assert np.allclose(fdr, fd[0, 1:], rtol=1e-6, atol=0)
