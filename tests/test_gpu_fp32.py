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
