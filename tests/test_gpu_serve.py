"""Serving engine (include/wbfmm_serve.h) on the GPU.

Served solves against the oracle FMM on the same tree (the engine's domain
and depth are reproduced here) at rounding level, the engine's error report
against the oracle's FP64 direct sum, identical bits for a request run
alone and inside a mix of concurrent requests, admission under a device
memory budget, device-resident arrays, concurrent submitting threads, the
autotuner, resident Helmholtz tables and the refusal of invalid requests."""

import threading

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

PARITY = 1e-10
LAP, HEL = 1, 2


def c2(z):
    return np.ascontiguousarray(np.stack([np.real(z), np.imag(z)], 1))


def lap_case(seed, n, nt=None, dipoles=False):
    rng = np.random.default_rng(seed)
    xs = rng.uniform(0, 1, (n, 3))
    q = rng.normal(size=n)
    xt = None if nt is None else rng.uniform(0, 1, (nt, 3))
    nrm = d = None
    if dipoles:
        nrm = rng.normal(size=(n, 3))
        nrm /= np.linalg.norm(nrm, axis=1)[:, None]
        d = rng.normal(size=n)
    return dict(problem=LAP, xs=xs, q=q, xt=xt, normals=nrm, dip=d)


def hel_case(seed, n, k, sphere=False, dipoles=False):
    rng = np.random.default_rng(seed)
    if sphere:
        v = rng.normal(size=(n, 3))
        xs = v / np.linalg.norm(v, axis=1)[:, None]
    else:
        xs = rng.uniform(-1, 1, (n, 3))
    q = c2(rng.normal(size=n) + 1j * rng.normal(size=n))
    nrm = d = None
    if dipoles:
        nrm = xs / np.linalg.norm(xs, axis=1)[:, None] if sphere else rng.normal(size=(n, 3))
        nrm = np.ascontiguousarray(nrm / np.linalg.norm(nrm, axis=1)[:, None])
        d = c2(rng.normal(size=n) + 1j * rng.normal(size=n))
    return dict(problem=HEL, xs=xs, q=q, xt=None, normals=nrm, dip=d, k=k)


def request(S, case, f, **kw):
    args = dict(case)
    problem = args.pop("problem")
    return S.make_request(problem, args.pop("xs"), q=args.pop("q"), xt=args.pop("xt"),
                          normals=args.pop("normals"), dip=args.pop("dip"), f=f,
                          k=args.pop("k", 0.0), **kw)


def out_array(S, case, field=3):
    n = len(case["xs"]) if case["xt"] is None else len(case["xt"])
    return np.zeros((n, S.field_width(case["problem"], field)))


def engine_domain(case):
    """the engine's tree domain: bounding cube widened by 1% of its width"""
    pts = case["xs"] if case["xt"] is None else np.concatenate([case["xs"], case["xt"]])
    lo, hi = pts.min(0), pts.max(0)
    e = float((hi - lo).max())
    return lo - 0.01 * e, 1.02 * e


@pytest.fixture(scope="module")
def S(wb):
    import wbfmm_amd.serve as S
    return S


def solve(S, srv, case, **kw):
    f = out_array(S, case, kw.get("field", 3))
    r = request(S, case, f, **kw)
    res = srv.wait(srv.submit(r, keep=f))
    assert res.status == 0, res.as_dict()
    return f, res


def test_serve_laplace_matches_oracle_fmm(S, oracle):
    case = lap_case(31, 3000, nt=900, dipoles=True)
    with S.Server(2) as srv:
        f, res = solve(S, srv, case, tol=1e-6, depth=3, check=300)
    assert res.depth == 3 and res.order == 13
    x0, D = engine_domain(case)
    g, _, _ = oracle.fmm(1, case["xs"], case["xt"], x0, D, 3, [13] * 17, [13] * 17, q=case["q"],
                         nrm=case["normals"], d=case["dip"], field=3)
    for c in range(4):
        assert np.abs(f[:, c] - g[:, c]).max() <= PARITY * np.abs(g[:, c]).max(), c
    # the reported error is the relative l2 error of every 3rd target
    # against the FP64 direct sum
    xt = case["xt"][::3][:300]
    fd = oracle.lap_direct(case["xs"], case["q"], xt, 3, case["normals"], case["dip"], 1e-6)
    fs = f[::3][:300]
    ep = np.linalg.norm(fs[:, 0] - fd[:, 0]) / np.linalg.norm(fd[:, 0])
    eg = np.linalg.norm(fs[:, 1:] - fd[:, 1:]) / np.linalg.norm(fd[:, 1:])
    assert abs(res.err_potential - ep) <= 1e-6 * ep + 1e-15
    assert abs(res.err_gradient - eg) <= 1e-6 * eg + 1e-15
    assert ep < 1e-5 and eg < 1e-4


def test_serve_helmholtz_matches_oracle_fmm(S, oracle):
    case = hel_case(32, 2500, 3.0, dipoles=True)
    with S.Server(2) as srv:
        f, res = solve(S, srv, case, order=10, depth=3, check=250)
    x0, D = engine_domain(case)
    # the oracle takes complex strengths as (n, 2) real arrays and returns
    # 8 doubles per target (potential, gradient; re and im)
    g, _, _ = oracle.fmm(2, case["xs"], case["xs"], x0, D, 3, [10] * 17, [10] * 17,
                         q=case["q"], nrm=case["normals"], d=case["dip"], k=3.0, field=3,
                         nthreads=4)
    for c in range(8):
        assert np.abs(f[:, c] - g[:, c]).max() <= PARITY * np.abs(g[:, c]).max(), c
    assert 0 < res.err_potential < 1e-3


def test_serve_helmholtz_tolerance_orders(S, oracle):
    # orders from the truncation number at tol 1e-6: the checked error
    # stays near the tolerance
    case = hel_case(33, 4000, 6.0, sphere=True, dipoles=True)
    with S.Server(2) as srv:
        f, res = solve(S, srv, case, tol=1e-6, check=200)
    assert res.order >= 20
    assert 0 < res.err_potential < 1e-5, res.as_dict()
    assert 0 < res.err_gradient < 1e-4, res.as_dict()


def mix_cases():
    return [
        lap_case(41, 20000, nt=5000),
        hel_case(42, 6000, 4.0),
        lap_case(43, 2000),
        hel_case(44, 3000, 8.0, sphere=True, dipoles=True),
        lap_case(45, 40000, dipoles=True),
        hel_case(46, 2000, 2.0),
        lap_case(47, 9000, nt=3000),
        lap_case(48, 1000),
    ]


def run_mix(S, srv, cases, kws):
    outs, tickets = [], []
    for c, kw in zip(cases, kws):
        f = out_array(S, c)
        outs.append(f)
        tickets.append(srv.submit(request(S, c, f, **kw), keep=f))
    res = [srv.wait(t) for t in tickets]
    for r in res:
        assert r.status == 0, r.as_dict()
    return outs, res


def test_serve_alone_equals_in_mix(S):
    cases = mix_cases()
    kws = [dict(tol=1e-6 if c["problem"] == LAP else 1e-4) for c in cases]
    with S.Server(4) as srv:
        alone = [run_mix(S, srv, [c], [kw])[0][0] for c, kw in zip(cases, kws)]
        mixed, res = run_mix(S, srv, cases, kws)
        st = srv.stats()
    for i, (a, m) in enumerate(zip(alone, mixed)):
        assert np.array_equal(a, m), i
    assert st.completed == 2 * len(cases)
    assert st.p50_ms <= st.p99_ms <= st.max_ms


def test_serve_budget_admission(S):
    cases = mix_cases()[:5]
    kws = [dict(tol=1e-6 if c["problem"] == LAP else 1e-4) for c in cases]
    with S.Server(4) as srv:
        fps = [srv.plan(request(S, c, out_array(S, c), **kw))[2] for c, kw in zip(cases, kws)]
        ref, _ = run_mix(S, srv, cases, kws)
    budget = max(fps) + 1
    with S.Server(4, budget) as srv:
        outs, res = run_mix(S, srv, cases, kws)
        st = srv.stats()
    # requests ran one at a time (each alone fits, no two fit together)
    assert st.peak_bytes <= budget
    for a, b in zip(ref, outs):
        assert np.array_equal(a, b)


def test_serve_device_arrays(S, wb):
    case = lap_case(51, 8000, nt=2000, dipoles=True)
    with S.Server(2) as srv:
        fh, _ = solve(S, srv, case, tol=1e-6)
        dev = {k: (wb.DeviceBuffer.from_array(v) if isinstance(v, np.ndarray) else v)
               for k, v in case.items()}
        fdev = wb.DeviceBuffer.from_array(np.zeros_like(fh))
        r = S.make_request(LAP, dev["xs"], q=dev["q"], xt=dev["xt"], normals=dev["normals"],
                           dip=dev["dip"], f=fdev, tol=1e-6, nsrc=8000, ntgt=2000)
        res = srv.wait(srv.submit(r, keep=(dev, fdev)))
        assert res.status == 0
        fd = fdev.download(np.zeros_like(fh))
        for b in list(dev.values()) + [fdev]:
            if isinstance(b, wb.DeviceBuffer):
                b.free()
    assert np.array_equal(fh, fd)


def test_serve_threads_submit_concurrently(S):
    cases = mix_cases()
    kws = [dict(tol=1e-6 if c["problem"] == LAP else 1e-4) for c in cases]
    with S.Server(3) as srv:
        ref, _ = run_mix(S, srv, cases, kws)
        got = [None] * len(cases)
        errs = []

        def client(idx):
            try:
                for i in idx:
                    o, _ = run_mix(S, srv, [cases[i]], [kws[i]])
                    got[i] = o[0]
            except Exception as e:  # reported below
                errs.append(e)

        th = [threading.Thread(target=client, args=(range(j, len(cases), 4),)) for j in range(4)]
        for t in th:
            t.start()
        for t in th:
            t.join()
    assert not errs, errs
    for a, b in zip(ref, got):
        assert np.array_equal(a, b)


def test_serve_autotuner(S):
    case = lap_case(61, 30000)
    with S.Server(2) as srv:
        r = request(S, case, out_array(S, case), tol=1e-4)
        best = srv.tune(r, [3, 4, 5])
        assert best in (3, 4, 5)
        assert srv.plan(r)[0] == best
        # a request of the same class and size band takes the tuned depth
        other = lap_case(62, 30000 + 500)
        assert srv.plan(request(S, other, out_array(S, other), tol=1e-4))[0] == best
        assert srv.stats().completed == 0  # tuning runs are not counted


def test_serve_autotuner_kernel_variants(S):
    # the autotuner also times the kernel variants at its best depth (the
    # sparse-level M2L tiles, the Helmholtz field kernel) and records them
    # for the class; a later request of the class runs with them and gives
    # the field of the library's own choice to within the precision used
    case = hel_case(93, 40000, 8.0, sphere=True, dipoles=True)
    with S.Server(2) as srv:
        f0, r0 = solve(S, srv, case, tol=1e-6, depth=6, check=300)
        r = request(S, case, out_array(S, case), tol=1e-6)
        best = srv.tune(r, [5, 6])
        d, vm, vf = srv.tuned(r)
        assert d == best and best in (5, 6)
        assert vm in (-1, 0, 2) and vf in (-1, 0, 1)
        f1, r1 = solve(S, srv, case, tol=1e-6, check=300)
        assert r1.depth == best
        assert 0 < r1.err_potential < 1e-5, r1.as_dict()
        if best == 6:
            rel = np.linalg.norm(f1 - f0) / np.linalg.norm(f0)
            assert rel < 1e-5, rel
    # a Laplace class: the M2L variants only (no field variant)
    lcase = lap_case(94, 30000)
    with S.Server(1) as srv:
        r = request(S, lcase, out_array(S, lcase), tol=1e-6)
        best = srv.tune(r, [4])
        d, vm, vf = srv.tuned(r)
        assert d == best == 4 and vm in (-1, 0, 1, 2) and vf == -1
        f, res = solve(S, srv, lcase, tol=1e-6, check=300)
        assert 0 < res.err_potential < 1e-5, res.as_dict()


def test_serve_rejects_invalid(S):
    case = lap_case(71, 100)
    with S.Server(1) as srv:
        f = out_array(S, case)
        bad = [dict(problem=7), dict(fstr=0), dict(order=99), dict(tol=0.0), dict(depth=11)]
        for b in bad:
            r = request(S, case, f, tol=1e-6)
            for k, v in b.items():
                setattr(r, k, v)
            with pytest.raises(RuntimeError):
                srv.submit(r, keep=f)
        # a valid request still runs afterwards
        f2, res = solve(S, srv, case, tol=1e-6)
        assert res.status == 0


def test_serve_fp32_path(S, oracle):
    # tol 1e-4 picks the FP32 path by default; with the same order the FP64
    # path differs by rounding only, and both meet the tolerance
    case = lap_case(81, 12000, nt=3000, dipoles=True)
    with S.Server(2) as srv:
        f32, r32 = solve(S, srv, case, tol=1e-4, precision=0, check=300)
        f64, r64 = solve(S, srv, case, tol=1e-4, precision=64, check=300)
        with pytest.raises(RuntimeError):   # no FP32 Laplace passes above order 15
            solve(S, srv, lap_case(82, 500), tol=1e-10, precision=32)
    assert r32.precision == 32 and r64.precision == 64 and r32.order == r64.order
    for c in range(4):
        assert np.abs(f32[:, c] - f64[:, c]).max() <= 3e-5 * np.abs(f64[:, c]).max(), c
    assert 0 < r32.err_potential < 1e-4 and 0 < r64.err_potential < 1e-4
    assert abs(r32.err_potential - r64.err_potential) < 0.1 * r64.err_potential


def test_serve_failed_request_reports_status_and_engine_survives(S):
    # the first device allocations of a request refused (forced from the
    # host, as when memory runs out): its result carries
    # WBFMM_HIP_ERR_NOMEM, its objects are released and the engine serves
    # the next requests, which give the same bits as a fresh engine
    cases = [lap_case(91, 6000), hel_case(92, 3000, 3.0)]
    kws = [dict(tol=1e-6), dict(tol=1e-4)]
    with S.Server(2) as srv:
        ref, _ = run_mix(S, srv, cases, kws)
    with S.Server(2) as srv:
        for nfail in (1, 3, 12):
            for c, kw in zip(cases, kws):
                f = out_array(S, c)
                srv.debug_fail_allocations(nfail)
                res = srv.wait(srv.submit(request(S, c, f, **kw), keep=f))
                assert res.status == -5, (nfail, res.as_dict())
        got, res = run_mix(S, srv, cases, kws)
        st = srv.stats()
    for a, b in zip(ref, got):
        assert np.array_equal(a, b)
    assert st.completed == len(cases)


def test_serve_helmholtz_tables_resident(S):
    # Helmholtz operator sets stay resident across requests (an engine-wide
    # cache keyed by k, domain, depth and orders): a problem solved again
    # after one of another wavenumber builds no tables, on whichever worker
    # takes it, and gives the same bits
    a, b = hel_case(95, 3000, 3.0), hel_case(96, 3000, 5.0)
    with S.Server(2) as srv:
        fa, ra = solve(S, srv, a, tol=1e-6)
        fb, rb = solve(S, srv, b, tol=1e-6)
        fa2, ra2 = solve(S, srv, a, tol=1e-6)
    assert ra.tables_ms > 0 and rb.tables_ms > 0
    assert ra2.tables_ms < 0.05 * ra.tables_ms, (ra.tables_ms, ra2.tables_ms)
    assert np.array_equal(fa, fa2)


def test_serve_refuses_request_above_budget(S):
    # a request whose estimated device memory exceeds the whole budget is
    # refused at submission; requests that fit still run
    small, big = lap_case(93, 2000), lap_case(94, 60000)
    with S.Server(2) as srv:
        fp_small = srv.plan(request(S, small, out_array(S, small), tol=1e-6))[2]
        fp_big = srv.plan(request(S, big, out_array(S, big), tol=1e-6))[2]
    assert fp_big > fp_small
    with S.Server(2, fp_small + 1) as srv:
        f = out_array(S, big)
        with pytest.raises(RuntimeError, match="-5"):
            srv.submit(request(S, big, f, tol=1e-6), keep=f)
        f, res = solve(S, srv, small, tol=1e-6)
        assert res.status == 0
