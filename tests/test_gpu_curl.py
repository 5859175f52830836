"""Laplace FMM with vector sources (nq components) on the GPU against the
oracle: the potentials and gradients of every component, the curl of the
vector potential and its gradient (src/local-laplace.c:632-1336,
src/direct-laplace.c:87-336).

The oracle runs one FMM per component (with second derivatives, field bit
4) and combines them into the curl (oracle.curl_from_components, pinned on
the CPU against the reference's closed-form curl).  The GPU tree holds the
nq components as planes of its level arrays and evaluates the curl from
the three components' local expansions plus the near-field curl kernel.
Parity bar: 1e-10 of each column's largest value.
"""

import ctypes

import numpy as np
import pytest

from conftest import fit_box

pytestmark = pytest.mark.gpu

PARITY = 1e-10


def colrel(a, b):
    return max(np.abs(a[:, c] - b[:, c]).max() / np.abs(b[:, c]).max() for c in range(b.shape[1]))


def solver(wb, xs, xt, depth, p, nq, host_mirror=True):
    x0, D = fit_box(xs, xt)
    S = wb.LaplaceSolver(xs, x0, D, depth, [p] * 17, [p] * 17, nq=nq, host_mirror=host_mirror)
    S.build_tree()
    return S, x0, D


def oracle_rows(oracle, xs, w, xt, x0, D, depth, p):
    return [oracle.fmm(1, xs, xt, x0, D, depth, [p] * 17, [p] * 17, q=np.ascontiguousarray(w[:, c]),
                       field=7)[0] for c in range(w.shape[1])]


@pytest.mark.parametrize("depth,p", [(3, 8), (4, 12)])
def test_curl_and_curl_gradient_match_oracle(wb, oracle, depth, p):
    rng = np.random.default_rng(900 + depth)
    n = 3000
    xs = rng.uniform(-1, 1, (n, 3))
    w = rng.normal(size=(n, 3))
    xt = np.concatenate([xs[::3], rng.uniform(-1, 1, (500, 3))])
    print(f"curl: n={n} depth={depth} p={p}", flush=True)
    S, x0, D = solver(wb, xs, xt, depth, p, 3)
    try:
        ws = np.ascontiguousarray(w)
        S.passes(ws, sstr=3)
        fc = np.zeros((len(xt), 3))
        S.field_points(xt, ws, fc, field=4, fstr=3, sstr=3)
        fg = np.zeros((len(xt), 12))
        S.field_points(xt, ws, fg, field=6, fstr=12, sstr=3)
        wb.sync()
    finally:
        S.close()
    want = oracle.curl_from_components(oracle_rows(oracle, xs, w, xt, x0, D, depth, p))
    assert colrel(fc, want[:, :3]) < PARITY
    assert colrel(fg, want) < PARITY


@pytest.mark.parametrize("nq", [2, 3, 5])
def test_vector_potentials_and_gradients_match_oracle(wb, oracle, nq):
    rng = np.random.default_rng(910 + nq)
    n = 2500
    xs = rng.uniform(0, 1, (n, 3))
    w = rng.normal(size=(n, nq))
    depth, p = 3, 10
    S, x0, D = solver(wb, xs, xs, depth, p, nq)
    try:
        ws = np.ascontiguousarray(w)
        S.passes(ws, sstr=nq)
        f = np.zeros((n, 4 * nq))
        S.field_points(xs, ws, f, field=3, fstr=4 * nq, sstr=nq)
        wb.sync()
        # the level coefficients of each component, in the reference layout
        got_m = S.coefficients(depth, 1)
    finally:
        S.close()
    rows = oracle_rows(oracle, xs, w, xs, x0, D, depth, p)
    for c in range(nq):
        assert colrel(f[:, c:c + 1], rows[c][:, 0:1]) < PARITY, c
        assert colrel(f[:, nq + 3 * c:nq + 3 * c + 3], rows[c][:, 1:4]) < PARITY, c
    for c in range(nq):
        _, _, dump = oracle.fmm(1, xs, xs, x0, D, depth, [p] * 17, [p] * 17,
                                q=np.ascontiguousarray(w[:, c]), field=1, dump_level=depth)
        ref = dump[1].reshape(8 ** depth, (p + 1) ** 2)
        assert np.abs(got_m[c] - ref).max() <= PARITY * np.abs(ref).max(), c


def test_box_field_curl_matches_batched(wb):
    """the reference's per-point wbfmm_laplace_box_field with the curl
    field types (4: curl, 6: curl and its gradient; 2: gradients at
    component stride fstr) equals the batched evaluation"""
    rng = np.random.default_rng(920)
    n = 1500
    xs = rng.uniform(-1, 1, (n, 3))
    w = np.ascontiguousarray(rng.normal(size=(n, 3)))
    xt = rng.uniform(-0.9, 0.9, (40, 3))
    depth, p = 3, 9
    S, _, _ = solver(wb, xs, xt, depth, p, 3)
    L = wb.lib()
    L.wbfmm_laplace_box_field.restype = ctypes.c_int
    try:
        S.passes(w, sstr=3)
        fg = np.zeros((len(xt), 12))
        S.field_points(xt, w, fg, field=6, fstr=12, sstr=3)
        fgr = np.zeros((len(xt), 9))
        S.field_points(xt, w, fgr, field=2, fstr=9, sstr=3)
        wb.sync()
        for i, x in enumerate(xt):
            x = np.ascontiguousarray(x)
            b = L.wbfmm_point_box(S.t, depth, wb.ptr(x))
            for fld, width, ref in ((4, 3, fg[i, :3]), (7, 12, fg[i])):
                out = np.zeros(width)
                assert L.wbfmm_laplace_box_field(S.t, depth, int(b), wb.ptr(w), 3, None, 0, fld, 1,
                                                 wb.ptr(x), wb.ptr(out), 1, None) == 0
                assert np.abs(out - ref).max() <= 1e-12 * np.abs(ref).max()
            out = np.zeros(3 * 4)  # component stride 4
            assert L.wbfmm_laplace_box_field(S.t, depth, int(b), wb.ptr(w), 3, None, 0, 2, 1,
                                             wb.ptr(x), wb.ptr(out), 4, None) == 0
            got = np.concatenate([out[4 * c:4 * c + 3] for c in range(3)])
            assert np.abs(got - fgr[i]).max() <= 1e-12 * np.abs(fgr[i]).max()
    finally:
        S.close()


def test_direct_curl_matches_reference_closed_form(wb, oracle):
    """wbfmm_laplace_field_direct with nq = 3 and the curl field types
    against the reference's closed forms (written out in the CPU suite),
    and with nq = 2 potentials against the oracle per component"""
    import test_oracle_laplace as T
    rng = np.random.default_rng(930)
    ns = 300
    xs = np.ascontiguousarray(rng.uniform(0, 1, (ns, 3)))
    w = np.ascontiguousarray(rng.normal(size=(ns, 3)))
    xt = rng.uniform(-0.5, 1.5, (12, 3))
    L = wb.lib()
    want = T.reference_curl(xs, w, xt, True)
    for i, x in enumerate(xt):
        x = np.ascontiguousarray(x)
        for fld, width in ((4, 3), (6, 12)):
            out = np.zeros(width)
            assert L.wbfmm_laplace_field_direct(wb.ptr(xs), 3, None, 0, ns, wb.ptr(w), 3, None, 0,
                                                3, fld, wb.ptr(x), wb.ptr(out), 1) == 0
            assert np.abs(out - want[i, :width]).max() <= 1e-12 * np.abs(want[i]).max()
    w2 = np.ascontiguousarray(w[:, :2])
    for i, x in enumerate(xt):
        x = np.ascontiguousarray(x)
        out = np.zeros(2 * 3)  # component stride 3
        assert L.wbfmm_laplace_field_direct(wb.ptr(xs), 3, None, 0, ns, wb.ptr(w2), 2, None, 0,
                                            2, 1, wb.ptr(x), wb.ptr(out), 3) == 0
        for c in range(2):
            ref = oracle.lap_direct(xs, np.ascontiguousarray(w2[:, c]), x[None, :], field=1)[0, 0]
            assert abs(out[3 * c] - ref) <= 1e-12 * abs(ref)


def test_curl_full_size_against_direct(wb):
    """2e5 vortex points at depth 5, order 12: the curl and its gradient
    against the GPU direct sums on sampled targets (rel. L2, FMM accuracy)"""
    rng = np.random.default_rng(940)
    n = 200_000
    xs = rng.uniform(0, 1, (n, 3))
    w = np.ascontiguousarray(rng.normal(size=(n, 3)))
    S, x0, D = solver(wb, xs, xs, 5, 12, 3, host_mirror=False)
    try:
        S.passes(w, sstr=3)
        idx = np.sort(rng.choice(n, 800, replace=False))
        xt = np.ascontiguousarray(xs[idx])
        fg = np.zeros((len(xt), 12))
        S.field_points(xt, w, fg, field=6, fstr=12, sstr=3)
        wb.sync()
    finally:
        S.close()
    L = wb.lib()
    L.wbfmm_hip_lap_direct_vec.restype = ctypes.c_int
    L.wbfmm_hip_lap_direct_vec.argtypes = [ctypes.c_void_p] * 4 + [ctypes.c_int, ctypes.c_uint,
                                                                   ctypes.c_void_p, ctypes.c_uint,
                                                                   ctypes.c_int, ctypes.c_double,
                                                                   ctypes.c_void_p, ctypes.c_int]
    s4 = wb.DeviceBuffer.from_array(np.concatenate([xs, np.zeros((n, 1))], 1))
    t4 = wb.DeviceBuffer.from_array(np.concatenate([xt, np.zeros((len(xt), 1))], 1))
    dw = wb.DeviceBuffer.from_array(w)
    fd = wb.DeviceBuffer(8 * 12 * len(xt))
    fd.zero()
    assert L.wbfmm_hip_lap_direct_vec(s4.ptr, dw.ptr, None, None, 3, n, t4.ptr, len(xt), 6, 1e-6,
                                      fd.ptr, 12) == 0
    ref = fd.download(np.zeros((len(xt), 12)))
    for b in (s4, t4, dw, fd):
        b.free()
    for name, cols in (("curl", slice(0, 3)), ("curl gradient", slice(3, 12))):
        e = np.linalg.norm(fg[:, cols] - ref[:, cols]) / np.linalg.norm(ref[:, cols])
        print(f"curl full size: {name} rel L2 error {e:.2e}", flush=True)
        assert e < 1e-5
