"""Laplace FMM on the GPU against the CPU oracle (same inputs, same tree,
same orders) and against FP64 direct sums at full size.

Both sides compute in FP64 and the FMM result is fixed by the tree and the
orders, so parity is at rounding level: the tolerance below is 1e-10 of
the largest value of each field component.  The full-size case (10^6
sources, depth 7, order 12) is checked through properties that do not
need the oracle: error against the GPU FP64 direct sum on sampled
targets, linearity in the strengths and run-to-run determinism.
"""

import subprocess
import sys

import numpy as np
import pytest

from conftest import fit_box

pytestmark = pytest.mark.gpu

PARITY = 1e-10


def gpu_fmm(W, xs, q, xt, x0, D, depth, p, field=3, nrm=None, d=None, keep=False):
    S = W.LaplaceSolver(np.ascontiguousarray(xs), x0, D, depth, [p] * 17, [p] * 17,
                        targets=np.ascontiguousarray(xt),
                        normals=None if nrm is None else np.ascontiguousarray(nrm))
    width = (1 if field & 1 else 0) + (3 if field & 2 else 0)
    f = np.zeros((len(xt), 4))
    S.solve(np.ascontiguousarray(q), f, field=field, fstr=4,
            d=None if d is None else np.ascontiguousarray(d))
    W.sync()
    if keep:
        return f, S
    S.close()
    return f


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def cases():
    rng = np.random.default_rng(11)
    out = {}
    xs = rng.uniform(0, 1, (2000, 3))
    out["cube_d3_p8"] = (xs, rng.normal(size=2000), xs, 3, 8, None, None)
    xs = rng.uniform(-1, 1, (4000, 3))
    out["cube_d4_p12"] = (xs, rng.normal(size=4000), xs, 4, 12, None, None)
    xs = rng.uniform(-1, 1, (1500, 3))
    xt = rng.uniform(-1, 1, (900, 3))
    out["separate_targets_d3_p10"] = (xs, rng.normal(size=1500), xt, 3, 10, None, None)
    v = rng.normal(size=(3000, 3))
    xs = v / np.linalg.norm(v, axis=1)[:, None]
    out["sphere_d4_p8"] = (xs, rng.normal(size=3000), xs, 4, 8, None, None)
    xs = rng.uniform(0, 1, (1200, 3))
    nrm = rng.normal(size=(1200, 3))
    nrm /= np.linalg.norm(nrm, axis=1)[:, None]
    out["dipoles_d3_p9"] = (xs, rng.normal(size=1200), xs, 3, 9, nrm, rng.normal(size=1200))
    return out


CASES = cases()


@pytest.mark.parametrize("name", sorted(CASES))
def test_fmm_matches_oracle(wb, oracle, name):
    xs, q, xt, depth, p, nrm, d = CASES[name]
    x0, D = fit_box(xs, xt)
    print(f"case {name}: n={len(xs)} depth={depth} p={p}", flush=True)
    fo, _, _ = oracle.fmm(1, xs, xt, x0, D, depth, [p] * 17, [p] * 17, q=q, nrm=nrm, d=d,
                          field=3 if d is None else 1)
    fg = gpu_fmm(wb, xs, q, xt, x0, D, depth, p, field=3 if d is None else 1, nrm=nrm, d=d)
    ncomp = 4 if d is None else 1
    for c in range(ncomp):
        assert rel(fg[:, c], fo[:, c]) < PARITY, (name, c, rel(fg[:, c], fo[:, c]))


def test_coefficients_match_oracle(wb, oracle):
    # mps/mpr in the reference's packing and normalisation, level by level
    rng = np.random.default_rng(12)
    xs = rng.uniform(0, 1, (2500, 3))
    q = rng.normal(size=2500)
    x0, D = fit_box(xs)
    depth, p = 4, 10
    fg, S = gpu_fmm(wb, xs, q, xs, x0, D, depth, p, keep=True)
    try:
        for lvl in (2, 3, 4):
            _, _, dumps = oracle.fmm(1, xs, xs, x0, D, depth, [p] * 17, [p] * 17, q=q,
                                     field=3, dump_level=lvl)
            C = (p + 1) ** 2
            mps_o = dumps[0].reshape(-1, C)
            mpr_o = dumps[1].reshape(-1, C)
            mps_g = S.coefficients(lvl, 0)
            mpr_g = S.coefficients(lvl, 1)
            assert rel(mps_g, mps_o) < PARITY, (lvl, "mps")
            # the GPU skips boxes holding neither sources nor targets
            occ = np.abs(mps_o).sum(1) > 0
            assert rel(mpr_g[occ], mpr_o[occ]) < PARITY, (lvl, "mpr")
    finally:
        S.close()


def test_direct_sum_matches_oracle(wb, oracle):
    import ctypes
    rng = np.random.default_rng(13)
    xs = rng.uniform(-1, 1, (700, 3))
    q = rng.normal(size=700)
    nrm = rng.normal(size=(700, 3))
    nrm /= np.linalg.norm(nrm, axis=1)[:, None]
    d = rng.normal(size=700)
    L = wb.lib()
    for xf in rng.uniform(-1, 1, (5, 3)):
        xf = np.ascontiguousarray(xf)
        want = oracle.lap_direct(xs, q, xf[None], nrm=nrm, d=d)[0]
        f = np.zeros(3)
        L.wbfmm_laplace_field_direct(wb.ptr(np.ascontiguousarray(xs)), 3,
                                     wb.ptr(np.ascontiguousarray(nrm)), 3, 700, wb.ptr(q), 1,
                                     wb.ptr(d), 1, 1, 2, wb.ptr(xf), wb.ptr(f), 1)
        assert rel(f, want[1:]) < 1e-12
        f = np.zeros(1)
        L.wbfmm_laplace_field_direct(wb.ptr(np.ascontiguousarray(xs)), 3,
                                     wb.ptr(np.ascontiguousarray(nrm)), 3, 700, wb.ptr(q), 1,
                                     wb.ptr(d), 1, 1, 1, wb.ptr(xf), wb.ptr(f), 1)
        assert abs(f[0] - want[0]) < 1e-12 * abs(want[0])


def test_box_field_and_field_targets(wb, oracle):
    # the reference's per-point wbfmm_laplace_box_field and the batched
    # wbfmm_laplace_field_targets agree with the oracle
    rng = np.random.default_rng(14)
    xs = rng.uniform(0, 1, (1500, 3))
    q = rng.normal(size=1500)
    xt = rng.uniform(0, 1, (40, 3))
    x0, D = fit_box(xs, xt)
    depth, p = 3, 9
    fo, _, _ = oracle.fmm(1, xs, xt, x0, D, depth, [p] * 17, [p] * 17, q=q, field=3)
    S = wb.LaplaceSolver(xs, x0, D, depth, [p] * 17, [p] * 17)  # no target list: dense
    S.build_tree()
    S.passes(q)
    L = wb.lib()
    fb = np.zeros((len(xt), 4))
    for i, x in enumerate(xt):
        x = np.ascontiguousarray(x)
        b = L.wbfmm_point_box(S.t, depth, wb.ptr(x))
        pot = np.zeros(1)
        L.wbfmm_laplace_box_field(S.t, depth, b, wb.ptr(q), 1, None, 0, 1, 1, wb.ptr(x),
                                  wb.ptr(pot), 1, None)
        g = np.zeros(3)
        L.wbfmm_laplace_box_field(S.t, depth, b, wb.ptr(q), 1, None, 0, 3, 1, wb.ptr(x),
                                  wb.ptr(g), 3, None)
        fb[i, 0] = pot[0]
        fb[i, 1:] = g
    ft = np.zeros((len(xt), 4))
    rc = L.wbfmm_laplace_field_targets(S.t, wb.ptr(np.ascontiguousarray(xt)), 24, len(xt),
                                       wb.ptr(q), 1, None, 0, 3, wb.ptr(ft), 4)
    assert rc == 0
    S.close()
    for c in range(4):
        assert rel(fb[:, c], fo[:, c]) < PARITY
        assert rel(ft[:, c], fo[:, c]) < PARITY


def test_edge_cases(wb, oracle):
    rng = np.random.default_rng(15)
    # one source; targets spread over the domain (mostly empty boxes)
    xs = np.array([[0.3, 0.6, 0.2]])
    q = np.array([1.7])
    xt = rng.uniform(0, 1, (200, 3))
    x0, D = np.zeros(3) - 0.01, 1.02
    fo, _, _ = oracle.fmm(1, xs, xt, x0, D, 3, [8] * 17, [8] * 17, q=q, field=3)
    fg = gpu_fmm(wb, xs, q, xt, x0, D, 3, 8)
    assert rel(fg, fo) < PARITY
    # every source in one leaf of a deep tree
    xs = 0.5 + 1e-4 * rng.uniform(0, 1, (300, 3))
    q = rng.normal(size=300)
    xt = rng.uniform(0, 1, (100, 3))
    fo, _, _ = oracle.fmm(1, xs, xt, x0, D, 5, [6] * 17, [6] * 17, q=q, field=3)
    fg = gpu_fmm(wb, xs, q, xt, x0, D, 5, 6)
    assert rel(fg, fo) < PARITY


def test_high_order_against_direct(wb, oracle):
    # order 24 (beyond any configuration's Laplace order) converges to the
    # FP64 direct sum; this exercises the quad-precision rotation tables
    rng = np.random.default_rng(16)
    xs = rng.uniform(0, 1, (3000, 3))
    q = rng.normal(size=3000)
    x0, D = fit_box(xs)
    fg = gpu_fmm(wb, xs, q, xs, x0, D, 3, 24)
    fd = oracle.lap_direct(xs, q, xs, cutoff=1e-6)
    for c in range(4):
        e = np.linalg.norm(fg[:, c] - fd[:, c]) / np.linalg.norm(fd[:, c])
        # truncation error at p=24 is ~3e-10 (potential), ~2e-9 (gradient);
        # FP64-recursion rotation tables would leave ~1e-4 here
        assert e < (1e-9 if c == 0 else 1e-8), (c, e)


def test_full_size_properties(wb):
    # 10^6 uniform sources, depth 7, order 12, targets = sources (the
    # plain-Laplace drop-in configuration)
    import ctypes
    n = 1_000_000
    rng = np.random.default_rng(3)
    xs = rng.uniform(0, 1, (n, 3))
    q1 = rng.normal(size=n)
    q2 = rng.normal(size=n)
    x0, D = fit_box(xs)
    print("full size: building", flush=True)
    S = wb.LaplaceSolver(xs, x0, D, 7, [12] * 17, [12] * 17, targets=xs, host_mirror=False)
    try:
        S.build_tree()
        f1 = np.zeros((n, 4))
        S.passes(q1)
        S.field(q1, f1)
        f2 = np.zeros((n, 4))
        S.passes(q2)
        S.field(q2, f2)
        f12 = np.zeros((n, 4))
        q12 = q1 + 2.0 * q2
        S.passes(q12)
        S.field(q12, f12)
        f1b = np.zeros((n, 4))
        S.passes(q1)
        S.field(q1, f1b)
        wb.sync()
        print("full size: solved", flush=True)
        # determinism: bitwise identical reruns
        assert np.array_equal(f1, f1b)
        # linearity in the strengths
        assert rel(f12, f1 + 2.0 * f2) < 1e-12
        # accuracy against the FP64 direct sum on 2000 sampled targets
        idx = rng.choice(n, 2000, replace=False)
        xt = np.ascontiguousarray(xs[idx])
        import oracle as O
        fd = O.lap_direct(xs, q1, xt, cutoff=1e-6)
        for c in range(4):
            e = np.linalg.norm(f1[idx, c] - fd[:, c]) / np.linalg.norm(fd[:, c])
            print(f"full size: component {c} rel L2 error {e:.3e}", flush=True)
            # order 12's truncation error: 4.6e-7 (potential), up to 2e-6
            # per gradient component; order 14 meets 1e-6 on every one
            # (test_full_size_order14_meets_1e6_on_every_component)
            assert e < (1e-6 if c == 0 else 5e-6), (c, e)
    finally:
        S.close()


def test_full_size_order14_meets_1e6_on_every_component(wb):
    # the same 10^6-point tree at order 14, the lowest order whose
    # potential and gradient both reach the metric's 1e-6 (at order 12 the
    # gradient's truncation error is 2.3e-6): every component within 1e-6
    # of the FP64 direct sum on 2000 sampled targets
    n = 1_000_000
    rng = np.random.default_rng(3)
    xs = rng.uniform(0, 1, (n, 3))
    q1 = rng.normal(size=n)
    x0, D = fit_box(xs)
    print("full size order 14: building", flush=True)
    S = wb.LaplaceSolver(xs, x0, D, 7, [14] * 17, [14] * 17, targets=xs, host_mirror=False)
    try:
        S.build_tree()
        f1 = np.zeros((n, 4))
        S.passes(q1)
        S.field(q1, f1)
        wb.sync()
        idx = rng.choice(n, 2000, replace=False)
        xt = np.ascontiguousarray(xs[idx])
        import oracle as O
        fd = O.lap_direct(xs, q1, xt, cutoff=1e-6)
        for c in range(4):
            e = np.linalg.norm(f1[idx, c] - fd[:, c]) / np.linalg.norm(fd[:, c])
            print(f"full size order 14: component {c} rel L2 error {e:.3e}", flush=True)
            assert e < 1e-6, (c, e)
    finally:
        S.close()


def test_point_outside_domain_aborts(wb):
    # the reference aborts (g_error) on a point outside the tree's cube
    code = (
        "import numpy as np, wbfmm_amd as W;"
        "xs=np.array([[0.5,0.5,0.5],[1.5,0.5,0.5]]);"
        "S=W.LaplaceSolver(xs,np.zeros(3),1.0,2,[4]*17,[4]*17);S.build_tree()"
    )
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True,
                       cwd=__import__("os").path.dirname(__import__("os").path.dirname(__file__)),
                       timeout=120)
    assert r.returncode != 0
    assert "do not lie in box" in r.stderr


def test_mfma_f64_operand_layout(wb):
    # the M2L kernel assumes lane l supplies A[l%16][l/16], B[l/16][l%16]
    # and receives D[l/16 + 4q][l%16]; check it against numpy
    out = np.zeros(256)
    assert wb.lib().wbfmm_hip_mfma_layout_probe(wb.ptr(out)) == 0
    i = np.arange(16)[:, None]
    k = np.arange(4)[None, :]
    A = 1.0 + i + 100.0 * k
    Bm = 1.0 + 3.0 * np.arange(4)[:, None] + 1000.0 * np.arange(16)[None, :]
    Dm = A @ Bm
    got = np.zeros((16, 16))
    for l in range(64):
        for q in range(4):
            got[l // 16 + 4 * q, l % 16] = out[4 * l + q]
    assert np.array_equal(got, Dm)


@pytest.mark.parametrize("p", [2, 4, 7, 12, 13, 15, 16, 18, 20, 24, 25])
def test_m2l_mfma_orders_match_oracle(wb, oracle, p):
    # the matrix-core M2L over the orders it serves (2..24; above 15 its
    # blocks span several bins and the M2M / L2L use the vector kernels)
    # and the vector M2L above them (25)
    rng = np.random.default_rng(20 + p)
    xs = rng.uniform(0, 1, (1800, 3))
    q = rng.normal(size=1800)
    x0, D = fit_box(xs)
    fo, _, _ = oracle.fmm(1, xs, xs, x0, D, 3, [p] * 17, [p] * 17, q=q, field=3)
    fg = gpu_fmm(wb, xs, q, xs, x0, D, 3, p)
    for c in range(4):
        assert rel(fg[:, c], fo[:, c]) < PARITY, (p, c, rel(fg[:, c], fo[:, c]))


def test_packed_sparse_level_m2l(wb, oracle, monkeypatch):
    # at a level where about 40 % of the boxes are occupied the M2L packs,
    # per offset, the targets whose source exists into 16-column tiles;
    # forced off and on, the two kernels agree to rounding
    rng = np.random.default_rng(77)
    n = 200_000
    xs = rng.uniform(0, 1, (n, 3))
    q = rng.normal(size=n)
    x0, D = fit_box(xs)
    out = {}
    for mode in ("0", "1"):
        monkeypatch.setenv("WBFMM_M2L_PACKED", mode)
        out[mode] = gpu_fmm(wb, xs, q, xs, x0, D, 6, 10)
    monkeypatch.delenv("WBFMM_M2L_PACKED")
    for c in range(4):
        assert rel(out["1"][:, c], out["0"][:, c]) < 1e-12, c
    idx = np.arange(0, n, n // 150)
    fd = oracle.lap_direct(xs, q, xs[idx], 3, cutoff=1e-6)
    assert np.linalg.norm(out["1"][idx, 0] - fd[:, 0]) / np.linalg.norm(fd[:, 0]) < 1e-5


def test_packed_sparse_level_separate_targets_dipoles(wb, monkeypatch):
    # the sparse-level kernel over a target list (boxes holding sources or
    # targets) with normal dipoles: forced off and on, the same field to
    # rounding
    rng = np.random.default_rng(78)
    n, nt = 150_000, 60_000
    xs = rng.uniform(0, 1, (n, 3))
    xt = rng.uniform(0, 1, (nt, 3))
    q = rng.normal(size=n)
    nrm = rng.normal(size=(n, 3))
    nrm /= np.linalg.norm(nrm, axis=1)[:, None]
    d = rng.normal(size=n)
    x0, D = fit_box(xs, xt)
    out = {}
    for mode in ("0", "1"):
        monkeypatch.setenv("WBFMM_M2L_PACKED", mode)
        out[mode] = gpu_fmm(wb, xs, q, xt, x0, D, 6, 9, nrm=nrm, d=d)
    monkeypatch.delenv("WBFMM_M2L_PACKED")
    for c in range(4):
        assert rel(out["1"][:, c], out["0"][:, c]) < 1e-12, c


def packed_cases():
    rng = np.random.default_rng(79)
    out = {}
    # about 60 % of the level-4 boxes occupied
    xs = rng.uniform(-1, 1, (4000, 3))
    out["cube_d4"] = (xs, rng.normal(size=4000), xs, 4, None, None)
    # sparse: points on a sphere, most boxes empty
    v = rng.normal(size=(6000, 3))
    xs = v / np.linalg.norm(v, axis=1)[:, None]
    out["sphere_d5"] = (xs, rng.normal(size=6000), xs, 5, None, None)
    # separate targets (a target list) with normal dipoles
    xs = rng.uniform(0, 1, (2500, 3))
    xt = rng.uniform(0, 1, (1100, 3))
    nrm = rng.normal(size=(2500, 3))
    nrm /= np.linalg.norm(nrm, axis=1)[:, None]
    out["targets_dipoles_d4"] = (xs, rng.normal(size=2500), xt, 4, nrm, rng.normal(size=2500))
    return out


PACKED_CASES = packed_cases()


@pytest.mark.parametrize("waves", ["4", "8"])
@pytest.mark.parametrize("p", [2, 3, 5, 8, 10, 11, 12])
@pytest.mark.parametrize("name", sorted(PACKED_CASES))
def test_packed_m2l_matches_oracle(wb, oracle, monkeypatch, name, p, waves):
    # the sparse-level M2L (forced at every level, four or eight wavefronts
    # per 32 targets) against the oracle FMM on the same tree
    xs, q, xt, depth, nrm, d = PACKED_CASES[name]
    x0, D = fit_box(xs, xt)
    print(f"packed case {name} p={p} waves={waves}", flush=True)
    fo, _, _ = oracle.fmm(1, xs, xt, x0, D, depth, [p] * 17, [p] * 17, q=q, nrm=nrm, d=d,
                          field=3, nthreads=8)
    monkeypatch.setenv("WBFMM_M2L_PACKED", "1")
    monkeypatch.setenv("WBFMM_PACKED_WAVES", waves)
    fg = gpu_fmm(wb, xs, q, xt, x0, D, depth, p, field=3, nrm=nrm, d=d)
    for c in range(4):
        assert rel(fg[:, c], fo[:, c]) < PARITY, (name, p, waves, c, rel(fg[:, c], fo[:, c]))


@pytest.mark.parametrize("p", [13, 14, 15])
@pytest.mark.parametrize("name", sorted(PACKED_CASES))
def test_packed_m2l_orders_13_15_match_oracle(wb, oracle, monkeypatch, name, p):
    # orders 13..15: the class-packed sparse-level M2L with sixteen
    # wavefronts per 32 targets (forced at every level) against the oracle
    xs, q, xt, depth, nrm, d = PACKED_CASES[name]
    x0, D = fit_box(xs, xt)
    print(f"packed case {name} p={p}", flush=True)
    fo, _, _ = oracle.fmm(1, xs, xt, x0, D, depth, [p] * 17, [p] * 17, q=q, nrm=nrm, d=d,
                          field=3, nthreads=8)
    monkeypatch.setenv("WBFMM_M2L_PACKED", "1")
    fg = gpu_fmm(wb, xs, q, xt, x0, D, depth, p, field=3, nrm=nrm, d=d)
    for c in range(4):
        assert rel(fg[:, c], fo[:, c]) < PARITY, (name, p, c, rel(fg[:, c], fo[:, c]))


def test_m2l_variants_match_oracle(wb, oracle):
    # the calling thread's M2L variant (wbfmm_hip_set_variant, as the
    # serving engine's autotuner sets it) at the sparse level of a tree
    # where the sparse-level kernel is chosen (25,000 uniform points at
    # depth 5, as below): chunks, offset tiles and class tiles each against
    # the oracle on the same tree
    rng = np.random.default_rng(81)
    n = 25_000
    xs = rng.uniform(0, 1, (n, 3))
    q = rng.normal(size=n)
    x0, D = fit_box(xs)
    fo, _, _ = oracle.fmm(1, xs, xs, x0, D, 5, [12] * 17, [12] * 17, q=q, field=3, nthreads=8)
    L = wb.lib()
    try:
        for v in (0, 1, 2):
            assert L.wbfmm_hip_set_variant(0, v) == 0
            assert L.wbfmm_hip_get_variant(0) == v
            print(f"m2l variant {v}", flush=True)
            fg = gpu_fmm(wb, xs, q, xs, x0, D, 5, 12)
            for c in range(4):
                assert rel(fg[:, c], fo[:, c]) < PARITY, (v, c, rel(fg[:, c], fo[:, c]))
    finally:
        L.wbfmm_hip_set_variant(0, -1)
    assert L.wbfmm_hip_set_variant(0, 3) != 0  # unknown variants refused
    assert L.wbfmm_hip_set_variant(1, 2) != 0
    assert L.wbfmm_hip_set_variant(2, 0) != 0


def test_packed_m2l_chosen_by_size_matches_oracle(wb, oracle):
    # 25,000 uniform points at depth 5: level 5 has about 17,500 occupied
    # boxes (53 %) in more than 1024 chunks, so the sparse-level kernel is
    # chosen there without forcing; compared with the oracle on the same
    # tree, at the headline order
    rng = np.random.default_rng(80)
    n = 25_000
    xs = rng.uniform(0, 1, (n, 3))
    q = rng.normal(size=n)
    x0, D = fit_box(xs)
    # the kernel's selection rule, restated: fewer than 60 % of the level's
    # boxes hold targets and the 16-target chunks of the eight Morton child
    # classes number at least 1024
    g = np.floor((xs - x0) / D * 32).astype(np.int64)
    boxes = np.unique(g[:, 0] + 32 * g[:, 1] + 1024 * g[:, 2])
    cls = (boxes & 1) | (((boxes >> 5) & 1) << 1) | (((boxes >> 10) & 1) << 2)
    chunks = sum((np.count_nonzero(cls == c) + 15) // 16 for c in range(8))
    assert len(boxes) / 32 ** 3 < 0.6 and chunks >= 1024, (len(boxes), chunks)
    fg = gpu_fmm(wb, xs, q, xs, x0, D, 5, 12)
    fo, _, _ = oracle.fmm(1, xs, xs, x0, D, 5, [12] * 17, [12] * 17, q=q, field=3, nthreads=8)
    for c in range(4):
        assert rel(fg[:, c], fo[:, c]) < PARITY, (c, rel(fg[:, c], fo[:, c]))
