"""Benchmark: Laplace FMM in FP64 through the drop-in API on one MI355X.

Workload (the plain-Laplace GPU drop-in configuration): 10^6 sources
uniform in the unit cube (SplitMix64 stream, seed 3), real charges iid
N(0,1), octree depth 7, expansion order 12 at every level, targets = the
source points, potential plus gradient.  A step is one whole solve with the
inputs already resident in HBM: tree build (Morton codes, sort, box
ranges, coefficient arrays, target list), P2M, upward pass (M2M), downward
pass (M2L + L2L) and field evaluation (L2P + near field), in the reference
driver's call order (tools/wbfmm-laplace-tree-calc.c.in:380-460).

Prints one JSON line: targets/s, per-step latency p50/p99, the relative
error against an FP64 direct sum on sampled targets, the roofline entry of
the dominant kernel (M2L) measured with HIP events on the library stream,
and a CPU baseline from the oracle on a bounded sample of the same
workload.
"""

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP64_PEAK_TFLOPS = 78.6   # MI355X dense FP64 (vector and matrix)
FP32_PEAK_TFLOPS = 157.3  # MI355X dense FP32 (vector and matrix)
HBM_PEAK_GBS = 8000.0
# HBM bytes per lap_m2l launch (mean over the 6 launches of one solve) from
# rocprofv3 FETCH_SIZE + WRITE_SIZE scaled by a copy kernel of known size;
# measured with the commands in profiles/lap_m2l_laplace_1e6_depth7_p12.md
M2L_TRAFFIC_BYTES_PER_LAUNCH = 21.89e9


def splitmix64(seed, n):
    """SplitMix64 stream, vectorised: n uint64 values."""
    with np.errstate(over="ignore"):
        k = np.arange(1, n + 1, dtype=np.uint64)
        z = np.uint64(seed) + k * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def uniform(seed, n):
    return (splitmix64(seed, n) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def normal(seed, n):
    """Box-Muller on two SplitMix64 streams."""
    u1 = uniform(seed, n)
    u2 = uniform(seed + 0x5bd1e995, n)
    return np.sqrt(-2.0 * np.log1p(-u1)) * np.cos(2.0 * np.pi * u2)


def workload(n=1_000_000, seed=3):
    xs = uniform(seed, 3 * n).reshape(n, 3)
    q = normal(seed + 17, n)
    lo, hi = xs.min(0), xs.max(0)
    x0 = lo - 0.01
    D = float((hi - lo).max()) + 0.02
    return xs, q, x0, D


def m2l_pairs(xs, x0, D, depth, xt=None):
    """Occupied-target x occupied-source list-4 pairs per level (targets xt,
    the sources when None).  A target at grid (i,j,k) interacts with the
    sources in the 6x6x6 block of its parent's neighbours' children, near
    neighbours excluded."""
    pairs = {}
    for l in range(2, depth + 1):
        nb = 1 << l

        def grid(x):
            g = np.floor((x - x0) / D * nb).astype(np.int64)
            g = np.clip(g, 0, nb - 1)
            o = np.zeros((nb + 6, nb + 6, nb + 6), dtype=np.int32)
            o[g[:, 0] + 3, g[:, 1] + 3, g[:, 2] + 3] = 1
            return o
        occ = grid(xs)
        core = (occ if xt is None else grid(xt))[3:nb + 3, 3:nb + 3, 3:nb + 3]
        tot = 0
        for px in range(2):
            for py in range(2):
                for pz in range(2):
                    t = core[px::2, py::2, pz::2]
                    for dx in range(-2 - px, 4 - px):
                        for dy in range(-2 - py, 4 - py):
                            for dz in range(-2 - pz, 4 - pz):
                                if abs(dx) <= 1 and abs(dy) <= 1 and abs(dz) <= 1:
                                    continue
                                s = occ[3 + px + dx:3 + px + dx + nb:2,
                                        3 + py + dy:3 + py + dy + nb:2,
                                        3 + pz + dz:3 + pz + dz + nb:2]
                                tot += int(np.sum(t * s))
        pairs[l] = tot
    return pairs


def cpu_baseline(p):
    """Oracle (CPU restatement of the reference path, not the reference's
    own AVX build, which cannot be built here) on a bounded sample with the
    same leaf density: 1e6/8 sources, depth 6, same order, M2L threaded
    over the host cores like the reference's downward pass."""
    import oracle
    n = 1_000_000 // 8
    xs, q, x0, D = workload(n, seed=3)
    nth = max(1, min(16, os.cpu_count() or 1))
    t0 = time.perf_counter()
    f, ph, _ = oracle.fmm(1, xs, xs, x0, D, 6, [p] * 17, [p] * 17, q=q, field=3, nthreads=nth)
    el = time.perf_counter() - t0
    return {"value": n / el, "unit": "targets/s", "cores": nth, "kind": "port",
            "sample": f"estimate from the oracle port (C restatement of the reference FMM; "
                      f"the reference's -O3 AVX build is not buildable here): {n} uniform "
                      f"sources = targets, depth 6 (same leaf density as 1e6 at depth 7), "
                      f"order {p}, potential+gradient, {el:.1f} s wall, M2L on {nth} threads"}


def sphere_points(seed, n):
    """n points on the unit sphere: normalised Gaussian triples (Box-Muller
    on SplitMix64 streams)"""
    g = np.stack([normal(seed + 101 * c, n) for c in range(3)], 1)
    return g / np.linalg.norm(g, axis=1)[:, None]


def hel_workload(kind):
    """(xs, q, nrm, d, xt, k, depth, orders, x0, D, description)"""
    if kind == "helmholtz-cube":
        # 1e4 monopoles uniform in [-1,1]^3, complex N(0,1) strengths, k = 1,
        # depth 4, order ceil(k w) + 10 = 11, 1e4 separate uniform targets
        n = 10_000
        xs = 2.0 * uniform(1, 3 * n).reshape(n, 3) - 1.0
        xt = 2.0 * uniform(2, 3 * n).reshape(n, 3) - 1.0
        q = normal(11, n) + 1j * normal(12, n)
        k, depth, orders = 1.0, 4, [11] * 17
        nrm = d = None
        desc = ("helmholtz_fp64_cube: 1e4 monopoles uniform in [-1,1]^3, complex N(0,1), k=1, "
                "depth 4, order 11, 1e4 separate targets, potential+gradient")
    else:
        # 1e6 points on the unit sphere, half monopoles and half radial normal
        # dipoles (complex N(0,1)), k = 16, depth 7, orders from the
        # truncation number at tol 1e-6 (25, 23, 23, 23, 23, 23 at levels
        # 2..7), targets = the points
        n = 1_000_000
        xs = sphere_points(4, n)
        half = np.arange(n) % 2 == 0
        q = np.where(half, normal(13, n) + 1j * normal(14, n), 0.0)
        d = np.where(~half, normal(15, n) + 1j * normal(16, n), 0.0)
        nrm = xs.copy()
        xt = xs
        k, depth = 16.0, 7
        orders = [25, 25, 25, 23, 23, 23, 23, 23] + [23] * 9
        desc = ("helmholtz_sphere (mixed precision: FP32 at levels with k a <= 2, FP64 above; "
                "--precision 64 for all FP64): 1e6 points on the unit sphere, 50% monopoles / 50% radial "
                "dipoles, complex N(0,1), k=16, depth 7, orders 25,23,23,23,23,23 (levels 2..7), "
                "targets = sources, potential+gradient")
    lo = np.minimum(xs.min(0), xt.min(0))
    hi = np.maximum(xs.max(0), xt.max(0))
    x0 = lo - 0.01
    D = float((hi - lo).max()) + 0.02
    return xs, q, nrm, d, xt, k, depth, orders, x0, D, desc


def run_helmholtz(args):
    import wbfmm_amd as W
    L = W.lib()
    xs, q, nrm, d, xt, k, depth, orders, x0, D, desc = hel_workload(args.workload)
    n, nt = len(xs), len(xt)
    c2 = lambda z: np.ascontiguousarray(np.stack([z.real, z.imag], 1))
    print(f"bench: workload {args.workload} ready", file=sys.stderr, flush=True)
    d_xs = W.DeviceBuffer.from_array(np.ascontiguousarray(xs))
    d_xt = d_xs if xt is xs else W.DeviceBuffer.from_array(np.ascontiguousarray(xt))
    d_q = W.DeviceBuffer.from_array(c2(q))
    d_d = None if d is None else W.DeviceBuffer.from_array(c2(d))
    d_n = None if nrm is None else W.DeviceBuffer.from_array(np.ascontiguousarray(nrm))
    d_f = W.DeviceBuffer(nt * 64)
    # the sphere configuration is defined in mixed precision (FP32 at the
    # low-frequency levels); --precision 64 runs it all in FP64
    prec = 32 if (args.workload == "helmholtz-sphere" and args.precision != "64") or \
        args.precision == "32" else 64
    t0 = time.perf_counter()
    S = W.HelmholtzSolver(d_xs, x0, D, depth, orders, orders, k, normals=d_n, nsrc=n,
                          targets=d_xt, ntgt=nt, precision=prec)
    W.sync()
    t_tables = time.perf_counter() - t0

    def step():
        d_f.zero()
        S.build_tree()
        S.passes(d_q, d=d_d)
        S.field(d_xt, d_q, d_f, field=3, fstr=8, ntgt=nt, d=d_d)

    for _ in range(args.warmup):
        step()
    W.sync()
    W.Timing.reset()
    W.Timing.enable(True)
    lat = []
    t_all = time.perf_counter()
    for _ in range(args.steps):
        t1 = time.perf_counter()
        step()
        W.sync()
        lat.append(time.perf_counter() - t1)
    total = time.perf_counter() - t_all
    W.Timing.enable(False)
    kern = {}
    for name in ("hel_m2l", "hel_m2m", "hel_l2l", "hel_p2m", "hel_l2p_p2p", "tree_sort",
                 "hel_m2l_f32", "hel_m2m_f32", "hel_l2l_f32", "hel_p2m_f32", "hel_l2p_p2p_f32",
                 "hel_to_f32", "hel_to_f64"):
        cnt, ms = W.Timing.query(name)
        if cnt:
            kern[name] = {"launches_per_step": cnt / args.steps, "ms_per_step": ms / args.steps,
                          "avg_ms": ms / cnt}
    print(f"bench: timed {args.steps} steps, {1e3 * total / args.steps:.2f} ms/step",
          file=sys.stderr, flush=True)
    f = d_f.download(np.zeros((nt, 8)))
    rng = np.random.default_rng(5)
    idx = np.sort(rng.choice(nt, min(args.check_targets, nt), replace=False))
    d_s4 = W.DeviceBuffer(n * 32)
    L.wbfmm_hip_stage4(d_xs.ptr, 24, n, d_s4.ptr)
    d_n4 = None
    if d_n is not None:
        d_n4 = W.DeviceBuffer(n * 32)
        L.wbfmm_hip_stage4(d_n.ptr, 24, n, d_n4.ptr)
    d_t4 = W.DeviceBuffer(len(idx) * 32)
    L.wbfmm_hip_stage4(W.DeviceBuffer.from_array(np.ascontiguousarray(xt[idx])).ptr, 24,
                       len(idx), d_t4.ptr)
    d_fd = W.DeviceBuffer(len(idx) * 64)
    d_fd.zero()
    L.wbfmm_hip_hel_direct(d_s4.ptr, d_q.ptr, None if d_n4 is None else d_n4.ptr,
                           None if d_d is None else d_d.ptr, n, d_t4.ptr, len(idx), k, 3, 1e-6,
                           d_fd.ptr, 8)
    fd = d_fd.download(np.zeros((len(idx), 8)))
    fc = f[idx][:, 0::2] + 1j * f[idx][:, 1::2]
    fdc = fd[:, 0::2] + 1j * fd[:, 1::2]
    err_pot = float(np.linalg.norm(fc[:, 0] - fdc[:, 0]) / np.linalg.norm(fdc[:, 0]))
    err_grad = float(np.linalg.norm(fc[:, 1:] - fdc[:, 1:]) / np.linalg.norm(fdc[:, 1:]))
    # roofline of the M2L: the reference's operation count per translation
    # of order N (two rotations, real H on complex coefficients: 4 flops per
    # entry of the (2n+1)^2 blocks; the complex coaxial product: 8 flops per
    # entry of the order-m blocks), times the occupied list-4 pairs of each
    # level
    pairs = m2l_pairs(xs, x0, D, depth, None if xt is xs else xt)

    def hel_flops_pair(N):
        Rn = (N + 1) * (2 * N + 1) * (2 * N + 3) // 3
        Cn = (N + 1) ** 2 + 2 * sum((N + 1 - m) ** 2 for m in range(1, N + 1))
        return 2 * 4 * Rn + 8 * Cn
    m2l_flops_step = sum(pairs[l] * hel_flops_pair(orders[l]) for l in pairs)
    # mixed precision: the M2L time is that of both kernels; the peak is
    # weighted by the flops each precision does (FP32 matrix peak 157.3)
    m64 = kern.get("hel_m2l", {"ms_per_step": 0.0, "launches_per_step": 0})
    m32 = kern.get("hel_m2l_f32", {"ms_per_step": 0.0, "launches_per_step": 0})
    m2l = {"ms_per_step": m64["ms_per_step"] + m32["ms_per_step"],
           "launches_per_step": max(m64["launches_per_step"] + m32["launches_per_step"], 1)}
    ka_thr = float(os.environ.get("WBFMM_HEL_FP64_KA", "2.0"))
    f32_flops = sum(pairs[l] * hel_flops_pair(orders[l]) for l in pairs
                    if prec == 32 and k * 0.5 * np.sqrt(3.0) * D / 2 ** l <= ka_thr)
    peak = FP64_PEAK_TFLOPS + (FP32_PEAK_TFLOPS - FP64_PEAK_TFLOPS) * f32_flops / m2l_flops_step
    achieved = m2l_flops_step / (m2l["ms_per_step"] * 1e-3) / 1e12
    roofline = {"bound": "mfma", "achieved": achieved, "peak": peak,
                "unit": "TFLOP/s", "frac": achieved / peak, "traffic": None,
                "kernel": "hel_m2l" + ("+hel_m2l_f32" if prec == 32 else ""),
                "flops_per_launch": m2l_flops_step / m2l["launches_per_step"],
                "pairs_per_step": int(sum(pairs.values())),
                "flops_per_pair": {str(l): hel_flops_pair(orders[l]) for l in pairs}}
    lat_ms = np.array(lat) * 1e3
    out = {
        "metric": "targets/sec at 1e-6 rel err vs FP64 direct; p50/p99 solve latency",
        "value": nt * args.steps / total, "unit": "targets/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": 1e3 * total / args.steps,
        "p50_ms": float(np.percentile(lat_ms, 50)), "p99_ms": float(np.percentile(lat_ms, 99)),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32/f64 mixed" if prec == 32 else "f64",
        "data": "synthetic (SplitMix64 streams)",
        "config": {"workload": desc, "n_sources": n, "n_targets": nt, "depth": depth,
                   "k": k, "orders": orders[2:depth + 1], "parallelism": "single GPU"},
        "error_vs_fp64_direct": {"targets_checked": int(len(idx)), "rel_l2_potential": err_pot,
                                 "rel_l2_gradient": err_grad},
        "roofline": roofline,
        "tables_s": t_tables,
        "kernels": kern,
    }
    S.close()
    if args.calibrate:
        calibration_copy(W, L)
    print(json.dumps(out))


def c2(z):
    return np.ascontiguousarray(np.stack([np.real(z), np.imag(z)], 1))


def mixed_problem(seed, n, helm, sphere, k):
    """one problem of the batched / serving workloads: n points uniform in
    [-1,1]^3 or on the unit sphere, N(0,1) strengths (complex for
    Helmholtz), targets = sources"""
    xs = sphere_points(seed, n) if sphere else 2.0 * uniform(seed, 3 * n).reshape(n, 3) - 1.0
    q = c2(normal(seed + 1, n) + 1j * normal(seed + 2, n)) if helm else normal(seed + 1, n)
    return np.ascontiguousarray(xs), q


def batched_problems(count=256, seed=6):
    """256 problems: N = floor(10^U(3,5)), Laplace and Helmholtz alternating
    (k ~ U(1,20)), cube or sphere with probability 1/2, orders from tol
    1e-6; `depth` is the fixed formula max(2, ceil(log8(N/20))), used with
    --batch-depth formula (by default the engine picks the depth from N and
    the kernel)"""
    u = uniform(seed, 4 * count).reshape(count, 4)
    out = []
    for i in range(count):
        n = int(10.0 ** (3.0 + 2.0 * u[i, 0]))
        helm = i % 2 == 1
        sphere = u[i, 1] < 0.5
        k = 1.0 + 19.0 * u[i, 2] if helm else 0.0
        depth = max(2, int(np.ceil(np.log(n / 20.0) / np.log(8.0))))
        xs, q = mixed_problem(seed * 1000003 + 17 * i, n, helm, sphere, k)
        out.append(dict(n=n, helm=helm, sphere=sphere, k=k, depth=depth, tol=1e-6, xs=xs, q=q))
    return out


def serving_requests(count, qps, seed=7):
    """Poisson arrivals at qps: N = floor(10^U(3,6)), Laplace or Helmholtz
    (k ~ U(1,32)) with probability 1/2, cube or sphere, accuracy 1e-4 or
    1e-8 with probability 1/2"""
    u = uniform(seed, 6 * count).reshape(count, 6)
    t, out = 0.0, []
    for i in range(count):
        n = int(10.0 ** (3.0 + 3.0 * u[i, 0]))
        helm = u[i, 1] < 0.5
        k = 1.0 + 31.0 * u[i, 2] if helm else 0.0
        tol = 1e-4 if u[i, 3] < 0.5 else 1e-8
        sphere = u[i, 4] < 0.5
        t += -np.log1p(-u[i, 5]) / qps
        xs, q = mixed_problem(seed * 1000003 + 17 * i, n, helm, sphere, k)
        out.append(dict(n=n, helm=helm, sphere=sphere, k=k, tol=tol, t=t, xs=xs, q=q))
    return out


def _upload(W, probs):
    for p in probs:
        p["d_xs"] = W.DeviceBuffer.from_array(p["xs"])
        p["d_q"] = W.DeviceBuffer.from_array(p["q"])
        p["d_f"] = W.DeviceBuffer(p["n"] * 8 * 8)
    W.sync()


PRECISION = {"auto": 0, "32": 32, "64": 64}


def _request(S, p, field=3, depth=0, check=0, precision=0):
    fw = 8 if p["helm"] else 4
    return S.make_request(2 if p["helm"] else 1, p["d_xs"], q=p["d_q"], f=p["d_f"], fstr=fw,
                          k=p["k"], tol=p["tol"], field=field, depth=depth, check=check,
                          nsrc=p["n"], precision=precision)


def _release(probs):
    for p in probs:
        for key in ("d_xs", "d_q", "d_f"):
            p[key].free()


def run_batched(args):
    """Config of many independent problems solved as one batch through the
    serving engine: a step is the whole batch (every problem submitted at
    once, the engine runs args.streams of them at a time)."""
    import wbfmm_amd as W
    import wbfmm_amd.serve as S
    probs = batched_problems(args.batch)
    _upload(W, probs)
    ntot = sum(p["n"] for p in probs)
    print(f"bench: {len(probs)} problems ready, {ntot} targets", file=sys.stderr, flush=True)
    srv = S.Server(args.streams)
    tuned = {}
    tune_s = 0.0
    if args.batch_depth == "tuned":
        # one autotuner run per request class (kernel, log2 N; the engine's
        # class also holds the precision and order, the same here for every
        # problem of a kernel): depths around the engine's own, then the
        # kernel variants at the best depth; untimed, like the table builds
        t_tune = time.perf_counter()
        for p in probs:
            key = ("helmholtz" if p["helm"] else "laplace", int(np.floor(np.log2(p["n"]))))
            if key in tuned:
                continue
            r = _request(S, p, precision=PRECISION[args.precision])
            d0 = srv.plan(r)[0]
            best = srv.tune(r, sorted({max(2, d0 - 1), d0, min(10, d0 + 1)}))
            _, vm, vf = srv.tuned(r)
            tuned[key] = (best, vm, vf)
        tune_s = time.perf_counter() - t_tune
        print(f"bench: tuned {len(tuned)} classes in {tune_s:.1f} s", file=sys.stderr, flush=True)

    def step(check):
        tk = [srv.submit(_request(S, p, depth=p["depth"] if args.batch_depth == "formula" else 0,
                                  check=check,
                                  precision=PRECISION[args.precision])) for p in probs]
        return [srv.wait(t) for t in tk]

    for _ in range(args.warmup):
        step(0)
    print("bench: warmup done", file=sys.stderr, flush=True)
    srv.reset_stats()
    t0 = time.perf_counter()
    res = []
    for s_ in range(args.steps):
        res = step(0)
    total = time.perf_counter() - t0
    st = srv.stats()
    # the error of each problem against the FP64 direct sum, from one more
    # (untimed) pass over the same problems: a solve gives the same bits
    # every time, so these are the errors of the timed results
    timed = res
    if args.check_targets:
        res = step(args.check_targets)
    srv.close()
    _release(probs)
    for p, r in zip(probs, res):
        print(f"bench: problem n={p['n']} {'helm k=%.1f' % p['k'] if p['helm'] else 'lap'} "
              f"{'sphere' if p['sphere'] else 'cube'} fp{r.precision} depth={r.depth} "
              f"order={r.order} tables={r.tables_ms:.1f} solve={r.solve_ms:.1f}",
              file=sys.stderr)
    bad = [r.status for r in res if r.status]
    ep = np.array([r.err_potential for r in res])
    eg = np.array([r.err_gradient for r in res])
    solve = np.array([r.solve_ms for r in timed])
    tables = np.array([r.tables_ms for r in timed])
    out = {
        "metric": "targets/sec at 1e-6 rel err vs FP64 direct; p50/p99 solve latency",
        "value": ntot * args.steps / total, "unit": "targets/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": 1e3 * total / args.steps,
        "p50_ms": st.p50_ms, "p99_ms": st.p99_ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if all(r.precision == 64 for r in res) else "f32/f64 mixed",
        "data": "synthetic (SplitMix64 seed 6)",
        "config": {"workload": f"batched: {len(probs)} problems, N = 10^U(3,5), Laplace / "
                               "Helmholtz alternating (k ~ U(1,20)), cube or sphere, depth "
                               + ("max(2, ceil(log8(N/20)))" if args.batch_depth == "formula"
                                  else "and kernel variants from the engine's autotuner per "
                                       "class" if args.batch_depth == "tuned"
                                  else "from N by the engine (about 16 points per leaf for "
                                       "Laplace, 32 for Helmholtz)")
                               + ", orders and precision from tol 1e-6 (precision auto: "
                               "FP64 Laplace, mixed FP32/FP64 Helmholtz), targets = "
                               "sources, potential+gradient; one step = the whole batch",
                   "problems": len(probs), "targets_per_step": ntot, "streams": args.streams,
                   "parallelism": "single GPU"},
        "latency_note": "p50/p99 over problems: submission to field written, all problems "
                        "submitted at once; errors from an untimed pass over the same problems",
        "solve_ms_p50": float(np.percentile(solve, 50)),
        "solve_ms_p99": float(np.percentile(solve, 99)),
        "tables_ms_total": float(tables.sum()),
        **({"autotuner": {"seconds": tune_s,
                          "classes": {f"{k[0]} log2N={k[1]}": {"depth": v[0], "m2l_variant": v[1],
                                                               "field_variant": v[2]}
                                      for k, v in sorted(tuned.items())}}}
           if tuned else {}),
        "error_vs_fp64_direct": {"targets_checked_per_problem": args.check_targets,
                                 "rel_l2_potential_median": float(np.median(ep)),
                                 "rel_l2_potential_max": float(ep.max()),
                                 "rel_l2_gradient_median": float(np.median(eg)),
                                 "rel_l2_gradient_max": float(eg.max()),
                                 "problems_potential_le_1e-6": int((ep <= 1e-6).sum())},
        "failed": len(bad),
    }
    print(json.dumps(out))


def run_serving(args):
    """Open-loop serving: requests arrive as a Poisson stream at args.qps
    and are submitted at their arrival times; latency is submission to
    field written."""
    import threading
    import wbfmm_amd as W
    import wbfmm_amd.serve as S
    reqs = serving_requests(args.requests, args.qps)
    _upload(W, reqs)
    print(f"bench: {len(reqs)} requests ready", file=sys.stderr, flush=True)
    srv = S.Server(args.streams)
    # warm the workers' Laplace tables and the kernels with small requests
    warm = [p for p in reqs if p["n"] < 20000][:2 * args.streams]
    for t in [srv.submit(_request(S, p, precision=PRECISION[args.precision])) for p in warm]:
        srv.wait(t)
    srv.reset_stats()
    tickets = [None] * len(reqs)
    t0 = time.perf_counter()

    def feeder():
        for i, p in enumerate(reqs):
            delay = t0 + p["t"] - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            tickets[i] = srv.submit(_request(S, p, check=args.check_targets,
                                             precision=PRECISION[args.precision]))

    th = threading.Thread(target=feeder)
    th.start()
    th.join()
    res = [srv.wait(t) for t in tickets]
    total = time.perf_counter() - t0
    st = srv.stats()
    srv.close()
    _release(reqs)
    for p, r in zip(reqs, res):
        print(f"bench: request n={p['n']} {'helm k=%.1f' % p['k'] if p['helm'] else 'lap'} "
              f"tol={p['tol']:g} fp{r.precision} depth={r.depth} order={r.order} "
              f"queue={r.queue_ms:.1f} tables={r.tables_ms:.1f} solve={r.solve_ms:.1f} "
              f"err={r.err_potential:.2e}", file=sys.stderr)
    ntot = sum(p["n"] for p in reqs)
    lat = np.array([r.latency_ms for r in res])
    per_tol = {}
    for tol in (1e-4, 1e-8):
        sel = [i for i, p in enumerate(reqs) if p["tol"] == tol]
        if sel:
            e = np.array([max(res[i].err_potential, 0.0) for i in sel])
            per_tol[str(tol)] = {"requests": len(sel), "rel_l2_potential_max": float(e.max()),
                                 "rel_l2_potential_median": float(np.median(e)),
                                 "latency_p50_ms": float(np.percentile(lat[sel], 50)),
                                 "latency_p99_ms": float(np.percentile(lat[sel], 99))}
    out = {
        "metric": "targets/sec at 1e-6 rel err vs FP64 direct; p50/p99 solve latency",
        "value": ntot / total, "unit": "targets/s", "n_gpus": 1, "steps": 1, "warmup": 0,
        "ms_per_step": 1e3 * total, "p50_ms": st.p50_ms, "p99_ms": st.p99_ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (SplitMix64 seed 7)",
        "config": {"workload": f"serving: {len(reqs)} requests, Poisson arrivals at "
                               f"{args.qps} req/s, N = 10^U(3,6), Laplace or Helmholtz (k ~ "
                               "U(1,32)), cube or sphere, accuracy 1e-4 or 1e-8 (FP32 for Laplace at 1e-4 with --precision auto), "
                               "targets = sources, potential+gradient",
                   "requests": len(reqs), "qps": args.qps, "streams": args.streams,
                   "parallelism": "single GPU"},
        "latency_ms": {"p50": st.p50_ms, "p99": st.p99_ms, "mean": st.mean_ms, "max": st.max_ms},
        "by_accuracy": per_tol,
        "tables_ms_total": float(sum(r.tables_ms for r in res)),
        "precision_mix": {"fp32": sum(1 for r in res if r.precision == 32),
                          "fp64": sum(1 for r in res if r.precision == 64)},
        "failed": sum(1 for r in res if r.status),
    }
    print(json.dumps(out))


def lap_flops_pair(p):
    """flops of one Laplace M2L translation (two rotations, one coaxial
    shift) at order p: 2 x 2R(p) + 2C(p)"""
    R = (p + 1) * (2 * p + 1) * (2 * p + 3) // 3
    C = (p + 1) ** 2 + 2 * sum((p + 1 - m) ** 2 for m in range(1, p + 1))
    return 2 * (2 * R) + 2 * C


def direct_reference(W, L, xs, d_xs, d_q, n, idx):
    """FP64 direct sum (GPU) at the sampled targets, self-cutoff 1e-6"""
    d_src4 = W.DeviceBuffer(n * 32)
    L.wbfmm_hip_stage4(d_xs.ptr, 24, n, d_src4.ptr)
    d_t4 = W.DeviceBuffer(len(idx) * 32)
    L.wbfmm_hip_stage4(W.DeviceBuffer.from_array(np.ascontiguousarray(xs[idx])).ptr, 24,
                       len(idx), d_t4.ptr)
    d_fd = W.DeviceBuffer(len(idx) * 32)
    d_fd.zero()
    L.wbfmm_hip_lap_direct(d_src4.ptr, d_q.ptr, None, None, n, d_t4.ptr, len(idx), 3, 1e-6,
                           d_fd.ptr, 4)
    fd = d_fd.download(np.zeros((len(idx), 4)))
    d_src4.free()
    d_t4.free()
    d_fd.free()
    return fd


def field_errors(f, fd):
    """relative L2 errors of potential and gradient, and max|diff|/max|f|
    per component (tests/maxdiff of the reference)"""
    err_pot = float(np.linalg.norm(f[:, 0] - fd[:, 0]) / np.linalg.norm(fd[:, 0]))
    err_grad = float(np.linalg.norm(f[:, 1:] - fd[:, 1:]) / np.linalg.norm(fd[:, 1:]))
    maxd = [float(np.abs(f[:, c] - fd[:, c]).max() / np.abs(fd[:, c]).max()) for c in range(4)]
    return err_pot, err_grad, maxd


KERNEL_NAMES = ("lap_m2l", "lap_m2m", "lap_l2l", "lap_p2m", "lap_l2p_p2p", "tree_sort",
                "tree_morton", "tree_box_ranges", "tree_box_flags", "tree_box_scatter",
                "tree_mark_boxes", "gather4", "gather_strength", "stage4")


def laplace_timed(W, L, d_xs, d_q, d_f, x0, D, n, depth, p, steps, warmup):
    """warmup + timed solves of the Laplace drop-in: a step is tree build,
    P2M, upward pass, downward pass and field at every target, inputs in
    HBM; the first solve's extra time is the table precompute"""
    S = W.LaplaceSolver(d_xs, x0, D, depth, [p] * 17, [p] * 17, targets=d_xs,
                        ntgt=n, nsrc=n, host_mirror=False)

    def step():
        d_f.zero()
        S.build_tree()
        S.passes(d_q)
        S.field(d_q, d_f, field=3, fstr=4)

    W.sync()
    t0 = time.perf_counter()
    step()
    W.sync()
    first = time.perf_counter() - t0
    for _ in range(max(warmup - 1, 0)):
        step()
    W.sync()
    print(f"bench: order {p} warmup done", file=sys.stderr, flush=True)
    try:
        import torch
        torch_sync = torch.cuda.synchronize if torch.cuda.is_available() else (lambda: None)
    except Exception:
        torch_sync = lambda: None

    W.Timing.reset()
    W.Timing.enable(True)
    lat = []
    torch_sync()
    W.sync()
    t_all = time.perf_counter()
    for _ in range(steps):
        t0 = time.perf_counter()
        step()
        W.sync()
        lat.append(time.perf_counter() - t0)
    W.sync()
    torch_sync()
    total = time.perf_counter() - t_all
    W.Timing.enable(False)
    kern = {}
    for name in KERNEL_NAMES:
        cnt, ms = W.Timing.query(name)
        if cnt:
            kern[name] = {"launches_per_step": cnt / steps, "ms_per_step": ms / steps,
                          "avg_ms": ms / cnt}
    S.close()
    print(f"bench: order {p}: timed {steps} steps, {1e3 * total / steps:.2f} ms/step",
          file=sys.stderr, flush=True)
    return {"total": total, "lat": lat, "kernels": kern,
            "precompute_ms": 1e3 * max(first - total / steps, 0.0)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--depth", type=int, default=7)
    ap.add_argument("--order", type=int, default=12)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--check-targets", type=int, default=2000)
    ap.add_argument("--no-accuracy-line", action="store_true",
                    help="skip the extra run at the order that meets 1e-6 on every component")
    ap.add_argument("--workload", default="laplace",
                    choices=["laplace", "helmholtz-cube", "helmholtz-sphere", "batched",
                             "serving"],
                    help="laplace: the headline configuration (default); helmholtz-cube / "
                         "helmholtz-sphere: the Helmholtz configurations; batched: 256 "
                         "independent problems per step; serving: Poisson request stream")
    ap.add_argument("--streams", type=int, default=8, help="serving engine streams")
    ap.add_argument("--batch", type=int, default=256, help="problems per batched step")
    ap.add_argument("--batch-depth", default="engine", choices=["engine", "formula", "tuned"],
                    help="batched: depth picked by the engine from N, the fixed formula "
                         "max(2, ceil(log8(N/20))), or the engine's autotuner (depth and "
                         "kernel variants per request class, tuned before the timed steps)")
    ap.add_argument("--requests", type=int, default=48, help="serving requests")
    ap.add_argument("--qps", type=float, default=4.0, help="serving arrival rate")
    ap.add_argument("--precision", default="auto", choices=["auto", "32", "64"],
                    help="batched / serving: auto = FP32 for Laplace requests at tol >= 1e-5, "
                         "mixed FP32/FP64 for Helmholtz at tol >= 1e-6")
    ap.add_argument("--calibrate", action="store_true",
                    help="end with a copy kernel of known size (wbfmm_hip_stage4 over 8Mi rows: "
                         "192 MiB read, 256 MiB written) to scale HBM counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: the solver runs on one GPU; --gpus {args.gpus} is not supported "
              "(sharding a tree over several GPUs is not implemented)", file=sys.stderr)
        sys.exit(2)
    if args.workload == "batched":
        run_batched(args)
        return
    if args.workload == "serving":
        run_serving(args)
        return
    if args.workload != "laplace":
        run_helmholtz(args)
        return

    import wbfmm_amd as W
    L = W.lib()
    n, depth, p = args.n, args.depth, args.order
    xs, q, x0, D = workload(n)
    print(f"bench: workload ready n={n} depth={depth} p={p}", file=sys.stderr, flush=True)

    # inputs resident in HBM before the timed region
    d_xs = W.DeviceBuffer.from_array(xs)
    d_q = W.DeviceBuffer.from_array(q)
    d_f = W.DeviceBuffer(n * 4 * 8)
    rng = np.random.default_rng(99)
    idx = np.sort(rng.choice(n, args.check_targets, replace=False))
    fd = direct_reference(W, L, xs, d_xs, d_q, n, idx)

    run = laplace_timed(W, L, d_xs, d_q, d_f, x0, D, n, depth, p, args.steps, args.warmup)
    kern = run["kernels"]
    f = d_f.download(np.zeros((n, 4)))
    err_pot, err_grad, maxd = field_errors(f[idx], fd)

    # roofline of the dominant kernel
    pairs = m2l_pairs(xs, x0, D, depth)
    flops_pair = lap_flops_pair(p)
    m2l_flops_step = flops_pair * sum(pairs.values())
    m2l = kern.get("lap_m2l", {"ms_per_step": float("nan"), "launches_per_step": 1})
    achieved = m2l_flops_step / (m2l["ms_per_step"] * 1e-3) / 1e12
    roofline = {"bound": "mfma", "achieved": achieved, "peak": FP64_PEAK_TFLOPS,
                "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS,
                "traffic": M2L_TRAFFIC_BYTES_PER_LAUNCH if (n, depth, p) == (1_000_000, 7, 12) else None,
                "kernel": "lap_m2l", "flops_per_launch": m2l_flops_step / m2l["launches_per_step"],
                "pairs_per_step": int(sum(pairs.values())), "flops_per_pair": flops_pair}

    lat_ms = np.array(run["lat"]) * 1e3
    tol = 1e-6
    out = {
        "metric": "targets/sec at 1e-6 rel err vs FP64 direct; p50/p99 solve latency",
        "value": n * args.steps / run["total"],
        "unit": "targets/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * run["total"] / args.steps,
        "p50_ms": float(np.percentile(lat_ms, 50)),
        "p99_ms": float(np.percentile(lat_ms, 99)),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (SplitMix64 seed 3, uniform cube, N(0,1) charges)",
        "config": {"workload": f"laplace_fp64_dropin: 1e6 sources = targets, unit cube, "
                               f"depth {depth}, order {p}, potential+gradient",
                   "n_sources": n, "n_targets": n, "depth": depth, "order": p,
                   "parallelism": "single GPU"},
        "error_vs_fp64_direct": {"targets_checked": int(len(idx)), "rel_l2_potential": err_pot,
                                 "rel_l2_gradient": err_grad, "max_rel_per_component": maxd},
        "meets_1e-6": {"potential": err_pot <= tol, "gradient": err_grad <= tol},
        "roofline": roofline,
        "kernels": kern,
        "precompute_ms": run["precompute_ms"],
    }
    # the same workload at the lowest order whose potential and gradient
    # both reach the metric's accuracy (the configured order is kept as the
    # headline; this line says what the bar costs)
    if not (err_pot <= tol and err_grad <= tol) and not args.no_accuracy_line:
        for pa in range(p + 1, min(p + 5, 25)):
            ra = laplace_timed(W, L, d_xs, d_q, d_f, x0, D, n, depth, pa, args.steps,
                               args.warmup)
            fa = d_f.download(np.zeros((n, 4)))
            ea_p, ea_g, _ = field_errors(fa[idx], fd)
            print(f"bench: order {pa}: {1e3 * ra['total'] / args.steps:.2f} ms/step, "
                  f"errors {ea_p:.2e} {ea_g:.2e}", file=sys.stderr, flush=True)
            if ea_p <= tol and ea_g <= tol:
                out["at_1e-6_all_components"] = {
                    "order": pa, "value": n * args.steps / ra["total"], "unit": "targets/s",
                    "ms_per_step": 1e3 * ra["total"] / args.steps,
                    "rel_l2_potential": ea_p, "rel_l2_gradient": ea_g}
                break
    if not args.no_cpu_baseline:
        print("bench: cpu baseline", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(p)
    if args.calibrate:
        calibration_copy(W, L)
    print(json.dumps(out))


def calibration_copy(W, L):
    """a copy kernel of known size for scaling rocprofv3's FETCH_SIZE and
    WRITE_SIZE: wbfmm_hip_stage4 over 8Mi rows, 24 B read and 32 B written
    per row"""
    rows = 8 << 20
    d_a = W.DeviceBuffer(rows * 24)
    d_b = W.DeviceBuffer(rows * 32)
    d_a.zero()
    W.sync()
    L.wbfmm_hip_stage4(d_a.ptr, 24, rows, d_b.ptr)
    W.sync()
    d_a.free()
    d_b.free()
    print("bench: calibration copy done", file=sys.stderr, flush=True)


if __name__ == "__main__":
    main()
