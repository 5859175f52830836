"""Helmholtz FMM on the GPU against the CPU oracle (same inputs, same tree,
same orders: parity at rounding level, 1e-10 of the largest value of each
field component) and against FP64 direct sums (truncation error)."""

import numpy as np
import pytest

from conftest import fit_box

pytestmark = pytest.mark.gpu

PARITY = 1e-10


def c2(z):
    """complex array -> (n, 2) float64 (re, im)"""
    return np.ascontiguousarray(np.stack([z.real, z.imag], 1))


def gpu_hel(W, xs, q, xt, x0, D, depth, p, k, nrm=None, d=None, field=3, keep=False):
    S = W.HelmholtzSolver(np.ascontiguousarray(xs), x0, D, depth, [p] * 17, [p] * 17, k,
                          normals=None if nrm is None else np.ascontiguousarray(nrm))
    S.build_tree()
    S.passes(None if q is None else c2(q), d=None if d is None else c2(d))
    f = np.zeros((len(xt), 8))
    S.field(np.ascontiguousarray(xt), None if q is None else c2(q), f, field=field, fstr=8,
            d=None if d is None else c2(d))
    W.sync()
    if keep:
        return f, S
    S.close()
    return f


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def cases():
    rng = np.random.default_rng(21)
    out = {}
    xs = rng.uniform(-1, 1, (1500, 3))
    xt = rng.uniform(-1, 1, (500, 3))
    q = rng.normal(size=1500) + 1j * rng.normal(size=1500)
    out["separate_d3_p8_k1.5"] = (xs, q, xt, 3, 8, 1.5, None, None)
    xs = rng.uniform(0, 1, (2500, 3))
    q = rng.normal(size=2500) + 1j * rng.normal(size=2500)
    out["equal_d4_p10_k3"] = (xs, q, xs, 4, 10, 3.0, None, None)
    xs = rng.uniform(-1, 1, (1200, 3))
    nrm = rng.normal(size=(1200, 3))
    nrm /= np.linalg.norm(nrm, axis=1)[:, None]
    q = rng.normal(size=1200) + 1j * rng.normal(size=1200)
    d = rng.normal(size=1200) + 1j * rng.normal(size=1200)
    out["dipoles_d3_p9_k2"] = (xs, q, xs, 3, 9, 2.0, nrm, d)
    v = rng.normal(size=(2000, 3))
    xs = v / np.linalg.norm(v, axis=1)[:, None]
    d = rng.normal(size=2000) + 1j * rng.normal(size=2000)
    out["sphere_dipoles_only_d3_p12_k4"] = (xs, None, xs, 3, 12, 4.0, xs.copy(), d)
    return out


CASES = cases()


@pytest.mark.parametrize("name", sorted(CASES))
def test_hel_fmm_matches_oracle(wb, oracle, name):
    xs, q, xt, depth, p, k, nrm, d = CASES[name]
    print(f"case {name}", flush=True)
    x0, D = fit_box(xs, xt)
    fg = gpu_hel(wb, xs, q, xt, x0, D, depth, p, k, nrm=nrm, d=d)
    fo, _, _ = oracle.fmm(2, xs, xt, x0, D, depth, [p] * 17, [p] * 17,
                          q=None if q is None else c2(q), nrm=nrm,
                          d=None if d is None else c2(d), k=k, field=3, nthreads=4)
    for c in range(8):
        assert rel(fg[:, c], fo[:, c]) < PARITY, (name, c, rel(fg[:, c], fo[:, c]))


def test_hel_coefficients_match_oracle(wb, oracle):
    rng = np.random.default_rng(5)
    xs = rng.uniform(-1, 1, (900, 3))
    q = rng.normal(size=900) + 1j * rng.normal(size=900)
    x0, D = fit_box(xs)
    depth, p, k = 3, 7, 2.5
    f, S = gpu_hel(wb, xs, q, xs, x0, D, depth, p, k, keep=True)
    errs = {}
    for lv in (3, 2):
        _, _, dumps = oracle.fmm(2, xs, xs, x0, D, depth, [p] * 17, [p] * 17, q=c2(q), k=k,
                                 field=1, dump_level=lv)
        C = (p + 1) ** 2
        for which in (0, 1):
            got = S.coefficients(lv, which)
            o = dumps[which].reshape(8 ** lv, C, 2)
            want = o[..., 0] + 1j * o[..., 1]
            errs[(lv, which)] = np.abs(got - want).max() / np.abs(want).max()
    S.close()
    print("relative coefficient differences (level, mps/mpr):", errs, flush=True)
    assert all(e < 1e-11 for e in errs.values()), errs


def test_hel_direct_matches_oracle(wb, oracle):
    rng = np.random.default_rng(9)
    xs = rng.uniform(-1, 1, (700, 3))
    xt = rng.uniform(-1, 1, (300, 3))
    q = rng.normal(size=700) + 1j * rng.normal(size=700)
    nrm = rng.normal(size=(700, 3))
    nrm /= np.linalg.norm(nrm, axis=1)[:, None]
    d = rng.normal(size=700) + 1j * rng.normal(size=700)
    k = 2.2
    W = wb
    n = len(xs)
    d_src = W.DeviceBuffer(n * 32)
    W.lib().wbfmm_hip_stage4(W.ptr(np.ascontiguousarray(xs)), 24, n, d_src.ptr)
    d_nrm = W.DeviceBuffer(n * 32)
    W.lib().wbfmm_hip_stage4(W.ptr(np.ascontiguousarray(nrm)), 24, n, d_nrm.ptr)
    d_t = W.DeviceBuffer(len(xt) * 32)
    W.lib().wbfmm_hip_stage4(W.ptr(np.ascontiguousarray(xt)), 24, len(xt), d_t.ptr)
    d_q = W.DeviceBuffer.from_array(c2(q))
    d_d = W.DeviceBuffer.from_array(c2(d))
    d_f = W.DeviceBuffer(len(xt) * 64)
    d_f.zero()
    assert W.lib().wbfmm_hip_hel_direct(d_src.ptr, d_q.ptr, d_nrm.ptr, d_d.ptr, n, d_t.ptr,
                                        len(xt), k, 3, 0.0, d_f.ptr, 8) == 0
    f = d_f.download(np.zeros((len(xt), 8)))
    fo = oracle.hel_direct(k, xs, q, xt, nrm=nrm, d=d)
    fc = f[:, 0::2] + 1j * f[:, 1::2]
    assert np.abs(fc - fo).max() < 1e-12 * np.abs(fo).max()


def test_hel_total_field_api(wb, oracle):
    # wbfmm_total_field / _grad (src/direct.c:35-294) at one point
    rng = np.random.default_rng(4)
    xs = np.ascontiguousarray(rng.uniform(-1, 1, (200, 3)))
    q = rng.normal(size=200) + 1j * rng.normal(size=200)
    xf = np.array([0.3, -0.2, 1.7])
    k = 1.3
    f = np.zeros(2)
    g = np.zeros(6)
    L = wb.lib()
    qq = c2(q)  # kept alive across the calls that read it
    assert L.wbfmm_total_field(k, wb.ptr(xs), 3, wb.ptr(qq), 2, None, 0, None, 0, 1, 200,
                               wb.ptr(xf), wb.ptr(f), 2) == 0
    assert L.wbfmm_total_field_grad(k, wb.ptr(xs), 3, wb.ptr(qq), 2, None, 0, None, 0, 1,
                                    200, wb.ptr(xf), wb.ptr(g), 6) == 0
    fo = oracle.hel_direct(k, xs, q, xf[None])[0]
    assert abs(f[0] + 1j * f[1] - fo[0]) < 1e-12 * abs(fo[0])
    assert np.abs(g[0::2] + 1j * g[1::2] - fo[1:]).max() < 1e-12 * np.abs(fo[1:]).max()


def test_hel_box_local_field_api(wb, oracle):
    # wbfmm_tree_box_local_field at single points equals the batched field
    rng = np.random.default_rng(8)
    xs = rng.uniform(-1, 1, (800, 3))
    q = rng.normal(size=800) + 1j * rng.normal(size=800)
    x0, D = fit_box(xs)
    f, S = gpu_hel(wb, xs, q, xs[:5], x0, D, 3, 8, 1.1, keep=True)
    L = wb.lib()
    qq = c2(q)  # kept alive across the calls that read it
    for i in range(5):
        x = np.ascontiguousarray(xs[i])
        b = L.wbfmm_point_box(S.t, 3, wb.ptr(x))
        fp = np.zeros(2)
        fgr = np.zeros(6)
        L.wbfmm_tree_box_local_field(S.t, 3, b, 1.1, wb.ptr(x), wb.ptr(fp), 2, wb.ptr(qq), 2,
                                     None, 0, 1, 1, None)
        L.wbfmm_tree_box_local_field(S.t, 3, b, 1.1, wb.ptr(x), wb.ptr(fgr), 6, wb.ptr(qq), 2,
                                     None, 0, 1, 2, None)
        assert np.abs(fp - f[i, :2]).max() < 1e-12 * np.abs(f[i, :2]).max()
        assert np.abs(fgr - f[i, 2:]).max() < 1e-12 * np.abs(f[i, 2:]).max()
    S.close()


def test_hel_config1_against_direct(wb, oracle):
    # the CPU configuration: 1e4 monopoles in [-1,1]^3, N(0,1) complex
    # strengths, k = 1, depth 4, order ceil(k w) + 10 = 11, 1e4 separate
    # targets, potential + gradient; against the FP64 direct sum
    rng = np.random.default_rng(1)
    xs = rng.uniform(-1, 1, (10000, 3))
    q = rng.normal(size=10000) + 1j * rng.normal(size=10000)
    xt = np.random.default_rng(2).uniform(-1, 1, (10000, 3))
    x0, D = fit_box(xs, xt)
    f = gpu_hel(wb, xs, q, xt, x0, D, 4, 11, 1.0)
    idx = np.arange(0, 10000, 25)
    fd = oracle.hel_direct(1.0, xs, q, xt[idx], cutoff=1e-6)
    fc = f[idx][:, 0::2] + 1j * f[idx][:, 1::2]
    for c in range(4):
        e = np.linalg.norm(fc[:, c] - fd[:, c]) / np.linalg.norm(fd[:, c])
        print(f"config1 component {c}: rel L2 error {e:.3e}", flush=True)
        assert e < 1e-4, (c, e)


def test_hel_target_list_matches_oracle(wb, oracle):
    # the target-list path (src/targets.c) against the oracle FMM on the
    # same tree, and against the batched field
    import ctypes
    rng = np.random.default_rng(12)
    xs = rng.uniform(-1, 1, (1500, 3))
    xt = np.ascontiguousarray(rng.uniform(-1, 1, (300, 3)))
    q = rng.normal(size=1500) + 1j * rng.normal(size=1500)
    x0, D = fit_box(xs, xt)
    k = 1.7
    f, S = gpu_hel(wb, xs, q, xt, x0, D, 3, 9, k, keep=True)
    L = wb.lib()
    tl = L.wbfmm_target_list_new(S.t, len(xt))
    L.wbfmm_target_list_add_points(tl, wb.ptr(xt), 24, len(xt))
    L.wbfmm_target_list_coefficients_init(tl, 3)
    L.wbfmm_target_list_local_coefficients.argtypes = [ctypes.c_void_p, ctypes.c_double,
                                                       ctypes.c_void_p]
    L.wbfmm_target_list_local_coefficients(tl, k, None)
    g = np.zeros((len(xt), 8))
    qq = c2(q)  # kept alive across the call that reads it
    L.wbfmm_target_list_local_field(tl, wb.ptr(qq), 2, None, 2, wb.ptr(g), 8)
    wb.sync()
    L.wbfmm_target_list_free(tl)
    S.close()
    fo, _, _ = oracle.fmm(2, xs, xt, x0, D, 3, [9] * 17, [9] * 17, q=qq, k=k, field=3,
                          nthreads=8)
    for c in range(8):
        assert rel(g[:, c], fo[:, c]) < PARITY, (c, rel(g[:, c], fo[:, c]))
    assert np.abs(g - f).max() < 1e-12 * np.abs(f).max()


@pytest.mark.parametrize("p", [2, 5, 13, 16, 20, 23, 24, 25, 27, 28])
def test_hel_m2l_mfma_orders_match_oracle(wb, oracle, p):
    # the matrix-core M2L, M2M and L2L over the orders they serve (2..27:
    # four wavefronts per block below 16, eight from 16) and the vector
    # kernels above (28); a second solve gives the same bits
    rng = np.random.default_rng(40 + p)
    xs = rng.uniform(-1, 1, (1500, 3))
    q = rng.normal(size=1500) + 1j * rng.normal(size=1500)
    k = 3.0
    print(f"order {p}", flush=True)
    x0, D = fit_box(xs)
    fg = gpu_hel(wb, xs, q, xs, x0, D, 3, p, k)
    fo, _, _ = oracle.fmm(2, xs, xs, x0, D, 3, [p] * 17, [p] * 17, q=c2(q), k=k, field=3,
                          nthreads=8)
    for c in range(8):
        assert rel(fg[:, c], fo[:, c]) < PARITY, (p, c, rel(fg[:, c], fo[:, c]))
    assert np.array_equal(fg, gpu_hel(wb, xs, q, xs, x0, D, 3, p, k))
