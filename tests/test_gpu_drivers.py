"""The replacement driver programs (wbfmm_amd/bin/wbfmm-laplace-tree-calc,
wbfmm_amd/bin/wbfmm-tree-calc: C programs built against include/wbfmm.h
that run the reference drivers' call sequence) and wbfmm_tree_sort_points.

The drivers read the reference's text point formats (M, MN, N, F) and write
"x y z f..." lines; their output is compared with the oracle FMM on the same
tree (domain fitted as with -b: bounding cube of sources and field points
widened by 0.01 on each side) at rounding level.  sort_points is checked
against Morton codes computed by the oracle: the records end up a stable
permutation in code order, for host and device arrays.
"""

import ctypes
import os
import subprocess

import numpy as np
import pytest

from conftest import fit_box

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "wbfmm_amd", "bin")
PARITY = 1e-10


def write_points(path, code, nq, rows):
    with open(path, "w") as fh:
        fh.write(f"{len(rows)} {nq} {code}\n")
        for r in rows:
            fh.write(" ".join(f"{v:.17e}" for v in r) + "\n")


def run_driver(name, args):
    exe = os.path.join(BIN, name)
    assert os.path.exists(exe), f"{exe} not built"
    r = subprocess.run([exe, *args], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    return np.array([[float(v) for v in line.split()] for line in r.stdout.splitlines()]), r.stderr


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("opts", ["", "-l", "-p", "-g", "-g -l", "-B"])
def test_laplace_driver_matches_oracle(oracle, tmp_path, opts):
    rng = np.random.default_rng(601)
    xs = rng.uniform(-1, 1, (2500, 3))
    q = rng.normal(size=2500)
    xf = rng.uniform(-1, 1, (700, 3))
    write_points(tmp_path / "src.dat", "M", 1, np.column_stack([xs, q]))
    write_points(tmp_path / "fld.dat", "F", 0, xf)
    out, err = run_driver("wbfmm-laplace-tree-calc",
                          ["-s", str(tmp_path / "src.dat"), "-f", str(tmp_path / "fld.dat"),
                           "-d", "3", "-S", "10", "-R", "10", "-b", *opts.split()])
    assert "downward pass completed" in err
    assert np.array_equal(out[:, :3], xf)
    x0, D = fit_box(xs, xf)
    fo, _, _ = oracle.fmm(1, xs, xf, x0, D, 3, [10] * 17, [10] * 17, q=q, field=3)
    want = fo[:, 1:] if "-g" in opts else fo[:, :1]
    assert out.shape[1] == 3 + want.shape[1]
    for c in range(want.shape[1]):
        assert rel(out[:, 3 + c], want[:, c]) < PARITY, (opts, c)


@pytest.mark.parametrize("code", ["MN", "N"])
def test_laplace_driver_dipole_formats(oracle, tmp_path, code):
    rng = np.random.default_rng(602)
    n = 1800
    xs = rng.uniform(0, 1, (n, 3))
    q = rng.normal(size=n)
    nrm = rng.normal(size=(n, 3))
    nrm /= np.linalg.norm(nrm, axis=1)[:, None]
    d = rng.normal(size=n)
    xf = rng.uniform(0, 1, (400, 3))
    rows = np.column_stack([xs, q, nrm, d]) if code == "MN" else np.column_stack([xs, nrm, d])
    write_points(tmp_path / "src.dat", code, 1, rows)
    write_points(tmp_path / "fld.dat", "F", 0, xf)
    out, _ = run_driver("wbfmm-laplace-tree-calc",
                        ["-s", str(tmp_path / "src.dat"), "-f", str(tmp_path / "fld.dat"),
                         "-d", "3", "-S", "9", "-R", "9", "-b"])
    x0, D = fit_box(xs, xf)
    fo, _, _ = oracle.fmm(1, xs, xf, x0, D, 3, [9] * 17, [9] * 17,
                          q=q if code == "MN" else None, nrm=nrm, d=d, field=1)
    assert rel(out[:, 3], fo[:, 0]) < PARITY


def test_laplace_driver_default_orders_and_origin(oracle, tmp_path):
    # no -S/-R: order 8 at every level; explicit origin and width
    rng = np.random.default_rng(603)
    xs = rng.uniform(0.1, 0.9, (1500, 3))
    q = rng.normal(size=1500)
    xf = rng.uniform(0.1, 0.9, (300, 3))
    write_points(tmp_path / "src.dat", "M", 1, np.column_stack([xs, q]))
    write_points(tmp_path / "fld.dat", "F", 0, xf)
    out, err = run_driver("wbfmm-laplace-tree-calc",
                          ["-s", str(tmp_path / "src.dat"), "-f", str(tmp_path / "fld.dat"),
                           "-d", "3", "-O", "0,0,0", "-D", "1"])
    assert "singular 8; regular 8" in err
    fo, _, _ = oracle.fmm(1, xs, xf, np.zeros(3), 1.0, 3, [8] * 17, [8] * 17, q=q, field=1)
    assert rel(out[:, 3], fo[:, 0]) < PARITY


@pytest.mark.parametrize("grad", [False, True])
def test_helmholtz_driver_matches_oracle(oracle, tmp_path, grad):
    rng = np.random.default_rng(604)
    xs = rng.uniform(-1, 1, (2000, 3))
    q = rng.normal(size=(2000, 2))
    xf = rng.uniform(-1, 1, (500, 3))
    k = 2.5
    write_points(tmp_path / "src.dat", "M", 2, np.column_stack([xs, q]))
    write_points(tmp_path / "fld.dat", "F", 0, xf)
    out, _ = run_driver("wbfmm-tree-calc",
                        ["-s", str(tmp_path / "src.dat"), "-f", str(tmp_path / "fld.dat"),
                         "-d", "3", "-S", "9", "-R", "9", "-k", str(k), "-b"] +
                        (["-g"] if grad else []))
    x0, D = fit_box(xs, xf)
    fo, _, _ = oracle.fmm(2, xs, xf, x0, D, 3, [9] * 17, [9] * 17, q=q, k=k, field=3,
                          nthreads=4)
    want = fo[:, 2:] if grad else fo[:, :2]
    assert out.shape[1] == 3 + want.shape[1]
    for c in range(want.shape[1]):
        assert rel(out[:, 3 + c], want[:, c]) < PARITY, (grad, c)


def test_helmholtz_driver_truncation_orders_and_dipoles(wb, oracle, tmp_path):
    # no -S/-R: orders from the truncation number of each level at -t;
    # MN format (monopoles, normals, dipole strengths)
    rng = np.random.default_rng(605)
    n = 1500
    v = rng.normal(size=(n, 3))
    xs = v / np.linalg.norm(v, axis=1)[:, None]
    q = rng.normal(size=(n, 2))
    d = rng.normal(size=(n, 2))
    xf = rng.uniform(-0.9, 0.9, (300, 3))
    k, tol = 4.0, 1e-5
    write_points(tmp_path / "src.dat", "MN", 2, np.column_stack([xs, q, xs, d]))
    write_points(tmp_path / "fld.dat", "F", 0, xf)
    out, err = run_driver("wbfmm-tree-calc",
                          ["-s", str(tmp_path / "src.dat"), "-f", str(tmp_path / "fld.dat"),
                           "-d", "3", "-k", str(k), "-t", str(tol), "-b"])
    x0, D = fit_box(xs, xf)
    orders = [0] + [oracle.lib().orc_truncation_number(k, D, l, tol) for l in range(1, 4)]
    for l in range(1, 4):
        assert f"level {l}: singular {orders[l]}; regular {orders[l]}" in err
    orders = orders + [orders[-1]] * (17 - len(orders))
    fo, _, _ = oracle.fmm(2, xs, xf, x0, D, 3, orders, orders, q=q, nrm=xs, d=d, k=k, field=1,
                          nthreads=4)
    for c in range(2):
        assert rel(out[:, 3 + c], fo[:, c]) < PARITY, c


@pytest.mark.parametrize("opts", ["", "-c", "-c -g"])
def test_laplace_driver_vector_sources(oracle, tmp_path, opts):
    """three-component sources (nq = 3): the potentials, the curl (-c) and
    the curl with its gradient (-c -g), widths as the reference driver's
    (tools/wbfmm-laplace-tree-calc.c.in:306-320)"""
    rng = np.random.default_rng(607)
    xs = rng.uniform(-1, 1, (2000, 3))
    w = rng.normal(size=(2000, 3))
    xf = rng.uniform(-1, 1, (400, 3))
    write_points(tmp_path / "src.dat", "M", 3, np.column_stack([xs, w]))
    write_points(tmp_path / "fld.dat", "F", 0, xf)
    out, _ = run_driver("wbfmm-laplace-tree-calc",
                        ["-s", str(tmp_path / "src.dat"), "-f", str(tmp_path / "fld.dat"),
                         "-d", "3", "-S", "9", "-R", "9", "-b", *opts.split()])
    x0, D = fit_box(xs, xf)
    rows = [oracle.fmm(1, xs, xf, x0, D, 3, [9] * 17, [9] * 17, q=np.ascontiguousarray(w[:, c]),
                       field=7)[0] for c in range(3)]
    if opts == "":
        want = np.column_stack([r[:, 0] for r in rows])
    else:
        want = oracle.curl_from_components(rows)[:, :3 if opts == "-c" else 12]
    assert out.shape[1] == 3 + want.shape[1]
    for c in range(want.shape[1]):
        assert rel(out[:, 3 + c], want[:, c]) < PARITY, (opts, c)


def test_drivers_refuse_unsupported_options(tmp_path):
    """the curl needs three source components; target lists take one"""
    write_points(tmp_path / "src.dat", "M", 1, np.array([[0.5, 0.5, 0.5, 1.0]]))
    write_points(tmp_path / "fld.dat", "F", 0, np.array([[0.2, 0.2, 0.2]]))
    r = subprocess.run([os.path.join(BIN, "wbfmm-laplace-tree-calc"), "-s",
                        str(tmp_path / "src.dat"), "-f", str(tmp_path / "fld.dat"), "-c"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode != 0 and "curl" in r.stderr
    write_points(tmp_path / "src3.dat", "M", 3, np.array([[0.5, 0.5, 0.5, 1.0, 2.0, 3.0]]))
    r = subprocess.run([os.path.join(BIN, "wbfmm-laplace-tree-calc"), "-s",
                        str(tmp_path / "src3.dat"), "-f", str(tmp_path / "fld.dat"), "-l"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode != 0 and "target list" in r.stderr


def morton_codes(oracle, pts, x0, D):
    L = oracle.lib()
    L.orc_point_index.restype = ctypes.c_uint64
    L.orc_point_index.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double]
    x0 = np.ascontiguousarray(x0, dtype=np.float64)
    out = []
    for p in pts:
        p = np.ascontiguousarray(p[:3], dtype=np.float64)
        out.append(L.orc_point_index(p.ctypes.data, x0.ctypes.data, D))
    return np.array(out, dtype=np.uint64)


@pytest.mark.parametrize("where", ["host", "device"])
def test_sort_points_stable_morton_order(wb, oracle, where):
    rng = np.random.default_rng(606)
    n = 3000
    # records: x y z and two payload columns; a coarse grid makes ties
    xyz = np.floor(rng.uniform(0, 1, (n, 3)) * 16) / 16 + 1e-3
    recs = np.ascontiguousarray(np.column_stack([xyz, np.arange(n), rng.normal(size=n)]))
    x0, D = np.zeros(3), 1.0
    L = wb.lib()
    L.wbfmm_tree_sort_points.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                         ctypes.c_uint]
    t = L.wbfmm_tree_new(x0.ctypes.data, D, n)
    try:
        if where == "host":
            got = recs.copy()
            assert L.wbfmm_tree_sort_points(t, got.ctypes.data, 40, n) == 0
        else:
            buf = wb.DeviceBuffer.from_array(recs)
            assert L.wbfmm_tree_sort_points(t, buf.ptr, 40, n) == 0
            got = buf.download(np.zeros_like(recs))
            buf.free()
    finally:
        L.wbfmm_tree_free(t)
    codes = morton_codes(oracle, recs, x0, D)
    order = np.argsort(codes, kind="stable")
    assert np.array_equal(got, recs[order])
