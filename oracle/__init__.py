"""ctypes bindings for the CPU oracle (test infrastructure only).

The oracle restates the reference's FMM evaluation path on the CPU
(oracle/wbfmm_oracle.c).  Only the test suite, the smoke check and the
CPU-baseline leg of bench.py load it, always as the checker and never as
the thing measured.  The product library never imports this package.
"""

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def build(force=False):
    """Compile the oracle with gcc (make -C oracle)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


class FmmCfg(ctypes.Structure):
    _fields_ = [
        ("problem", ctypes.c_int),
        ("k", ctypes.c_double),
        ("x0", ctypes.c_double * 3),
        ("D", ctypes.c_double),
        ("depth", ctypes.c_int),
        ("order_s", ctypes.c_int * 17),
        ("order_r", ctypes.c_int * 17),
        ("field", ctypes.c_int),
        ("nthreads", ctypes.c_int),
        ("dump_level", ctypes.c_int),
        ("dump_mps", _dp),
        ("dump_mpr", _dp),
        ("t_phase", ctypes.c_double * 6),
    ]


def lib():
    global _lib
    if _lib is None:
        build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_legendre.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, _dp]
        L.orc_bessel_j.argtypes = [ctypes.c_int, ctypes.c_double, _dp]
        L.orc_bessel_h.argtypes = [ctypes.c_int, ctypes.c_double, _dp]
        L.orc_truncation_number.argtypes = [ctypes.c_double, ctypes.c_double,
                                            ctypes.c_int, ctypes.c_double]
        L.orc_morton_encode.restype = ctypes.c_uint64
        L.orc_morton_encode.argtypes = [ctypes.c_uint32] * 3
        L.orc_point_index.restype = ctypes.c_uint64
        L.orc_point_index.argtypes = [_dp, _dp, ctypes.c_double]
        L.orc_box_list4.argtypes = [ctypes.c_int, ctypes.c_uint64,
                                    ctypes.POINTER(ctypes.c_uint64), _ip]
        L.orc_box_neighbours.argtypes = [ctypes.c_int, ctypes.c_uint64,
                                         ctypes.POINTER(ctypes.c_uint64)]
        L.orc_rotation_size.argtypes = [ctypes.c_int]
        L.orc_H_rotation.argtypes = [_dp, ctypes.c_int, ctypes.c_double]
        L.orc_shift_angles.argtypes = [ctypes.c_int] * 3 + [_dp] * 4
        L.orc_lap_p2m.argtypes = [ctypes.c_int, _dp, _dp, ctypes.c_double, _dp]
        L.orc_lap_p2m_normal.argtypes = [ctypes.c_int, _dp, _dp, _dp, ctypes.c_double, _dp]
        L.orc_lap_m_eval.restype = ctypes.c_double
        L.orc_lap_m_eval.argtypes = [ctypes.c_int, _dp, _dp, _dp]
        L.orc_lap_l_eval.argtypes = [ctypes.c_int, _dp, _dp, _dp, ctypes.c_int, _dp]
        L.orc_lap_l_hess.argtypes = [ctypes.c_int, _dp, _dp, _dp, _dp]
        L.orc_lap_coaxial_cfft.restype = ctypes.c_double
        L.orc_lap_coaxial_cfft.argtypes = [ctypes.c_int] * 4
        L.orc_lap_coaxial.argtypes = [ctypes.c_int, _dp, ctypes.c_int, _dp, ctypes.c_int,
                                      ctypes.c_double, ctypes.c_double]
        L.orc_lap_rotate.argtypes = [_dp, _dp, ctypes.c_int, _dp, ctypes.c_double,
                                     ctypes.c_double, ctypes.c_double]
        L.orc_lap_direct.argtypes = [ctypes.c_int, _dp, _dp, _dp, _dp, ctypes.c_int, _dp,
                                     ctypes.c_int, ctypes.c_double, _dp]
        L.orc_fmm.argtypes = [ctypes.POINTER(FmmCfg), ctypes.c_int, _dp, _dp, _dp, _dp,
                              ctypes.c_int, _dp, _dp]
        L.orc_hel_p2m.argtypes = [ctypes.c_double, ctypes.c_int, _dp, _dp, _dp, _dp]
        L.orc_hel_p2m_normal.argtypes = [ctypes.c_double, ctypes.c_int, _dp, _dp, _dp, _dp, _dp]
        L.orc_hel_h_eval.argtypes = [ctypes.c_double, ctypes.c_int, _dp, _dp, _dp, _dp]
        L.orc_hel_j_eval.argtypes = [ctypes.c_double, ctypes.c_int, _dp, _dp, _dp, ctypes.c_int, _dp]
        L.orc_hel_coaxial_table.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, _dp]
        L.orc_hel_coaxial.argtypes = [_dp, ctypes.c_int, _dp, ctypes.c_int, _dp, ctypes.c_int,
                                      ctypes.c_int, ctypes.c_double]
        L.orc_hel_rotate.argtypes = [_dp, _dp, ctypes.c_int, _dp, ctypes.c_double,
                                     ctypes.c_double, ctypes.c_double]
        L.orc_hel_direct.argtypes = [ctypes.c_double, ctypes.c_int, _dp, _dp, _dp, _dp,
                                     ctypes.c_int, _dp, ctypes.c_int, ctypes.c_double, _dp]
        _lib = L
    return _lib


def P(a):
    """numpy float64 array -> ctypes double pointer (None passes NULL)."""
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


def legendre(N, theta):
    out = np.zeros((N + 1) * (N + 2) // 2)
    lib().orc_legendre(N, np.cos(theta), np.sin(theta), P(out))
    return out


def rotation_size(N):
    return lib().orc_rotation_size(N)


def H_rotation(N, theta):
    H = np.zeros(rotation_size(N))
    lib().orc_H_rotation(P(H), N, theta)
    return H


def shift_angles(dx, dy, dz):
    v = [ctypes.c_double() for _ in range(4)]
    lib().orc_shift_angles(dx, dy, dz, *[ctypes.byref(x) for x in v])
    return tuple(x.value for x in v)


def lap_p2m(N, x0, xs, q, M=None):
    M = np.zeros((N + 1) ** 2) if M is None else M
    lib().orc_lap_p2m(N, P(np.ascontiguousarray(x0, float)),
                      P(np.ascontiguousarray(xs, float)), float(q), P(M))
    return M


def lap_p2m_normal(N, x0, xs, nrm, d, M=None):
    M = np.zeros((N + 1) ** 2) if M is None else M
    lib().orc_lap_p2m_normal(N, P(np.ascontiguousarray(x0, float)),
                             P(np.ascontiguousarray(xs, float)),
                             P(np.ascontiguousarray(nrm, float)), float(d), P(M))
    return M


def lap_m_eval(N, x0, M, xf):
    return lib().orc_lap_m_eval(N, P(np.ascontiguousarray(x0, float)), P(M),
                                P(np.ascontiguousarray(xf, float)))


def lap_l_eval(N, x0, L, xf, field=3):
    f = np.zeros(4)
    lib().orc_lap_l_eval(N, P(np.ascontiguousarray(x0, float)), P(L),
                         P(np.ascontiguousarray(xf, float)), field, P(f))
    return f


def lap_l_hess(N, x0, L, xf):
    """second derivatives (xx, yy, zz, xy, yz, zx) of the local expansion
    L (order N, about x0) at xf"""
    h = np.zeros(6)
    lib().orc_lap_l_hess(N, P(np.asarray(x0, float)), P(np.ascontiguousarray(L, float)),
                         P(np.asarray(xf, float)), P(h))
    return h


def lap_coaxial(kind, Ci, Ni, No, t):
    Co = np.zeros((No + 1) ** 2)
    lib().orc_lap_coaxial(kind, P(Co), No, P(np.ascontiguousarray(Ci, float)), Ni, t, 0.0)
    return Co


def lap_rotate(Ci, N, H, ph, ch):
    Co = np.zeros((N + 1) ** 2)
    lib().orc_lap_rotate(P(Co), P(np.ascontiguousarray(Ci, float)), N, P(H), ph, ch, 0.0)
    return Co


def lap_direct(xs, q, xt, field=3, nrm=None, d=None, cutoff=0.0):
    """rows: potential, gradient and, with field bit 4 (monopoles), the
    second derivatives (xx, yy, zz, xy, yz, zx)"""
    xs = np.ascontiguousarray(xs, float)
    xt = np.ascontiguousarray(xt, float)
    f = np.zeros((len(xt), 10 if field & 4 else 4))
    lib().orc_lap_direct(len(xs), P(xs), P(None if q is None else np.ascontiguousarray(q, float)),
                         P(None if nrm is None else np.ascontiguousarray(nrm, float)),
                         P(None if d is None else np.ascontiguousarray(d, float)),
                         len(xt), P(xt), field, cutoff, P(f))
    return f


def _c(a):
    return np.ascontiguousarray(a, float)


def curl_from_components(rows):
    """curl u = curl A and its gradient from the rows of the three
    components A_l (each [potential, gradient, xx, yy, zz, xy, yz, zx]):
    returns (n, 12): u, then d_j u_i at 3 + 3j + i, the layout of the
    reference's curl-gradient field (src/local-laplace.c:147-183)"""
    J = [r[:, 1:4] for r in rows]            # J[l][:, k] = d_k A_l
    H = [r[:, 4:10] for r in rows]
    hi = {(0, 0): 0, (1, 1): 1, (2, 2): 2, (0, 1): 3, (1, 0): 3, (1, 2): 4, (2, 1): 4,
          (0, 2): 5, (2, 0): 5}
    n = rows[0].shape[0]
    out = np.zeros((n, 12))
    out[:, 0] = J[2][:, 1] - J[1][:, 2]
    out[:, 1] = J[0][:, 2] - J[2][:, 0]
    out[:, 2] = J[1][:, 0] - J[0][:, 1]
    for j in range(3):
        out[:, 3 + 3 * j + 0] = H[2][:, hi[j, 1]] - H[1][:, hi[j, 2]]
        out[:, 3 + 3 * j + 1] = H[0][:, hi[j, 2]] - H[2][:, hi[j, 0]]
        out[:, 3 + 3 * j + 2] = H[1][:, hi[j, 0]] - H[0][:, hi[j, 1]]
    return out


def hel_p2m(k, N, x0, xs, q, M=None):
    """q complex scalar"""
    M = np.zeros(2 * (N + 1) ** 2) if M is None else M
    lib().orc_hel_p2m(k, N, P(_c(x0)), P(_c(xs)), P(_c([q.real, q.imag])), P(M))
    return M


def hel_p2m_normal(k, N, x0, xs, nrm, d, M=None):
    M = np.zeros(2 * (N + 1) ** 2) if M is None else M
    lib().orc_hel_p2m_normal(k, N, P(_c(x0)), P(_c(xs)), P(_c(nrm)), P(_c([d.real, d.imag])), P(M))
    return M


def hel_h_eval(k, N, x0, M, xf):
    f = np.zeros(2)
    lib().orc_hel_h_eval(k, N, P(_c(x0)), P(M), P(_c(xf)), P(f))
    return f[0] + 1j * f[1]


def hel_j_eval(k, N, x0, L, xf, field=3):
    f = np.zeros(8)
    lib().orc_hel_j_eval(k, N, P(_c(x0)), P(_c(L)), P(_c(xf)), field, P(f))
    return f[0::2] + 1j * f[1::2]


def coaxial_size(L):
    return (L + 1) * (L + 2) * (L + 3) // 6 + 8


def hel_coaxial_table(kind, L, kr):
    E = np.zeros(2 * ((L + 2) * (L + 3) * (L + 4) // 6))
    lib().orc_hel_coaxial_table(kind, L, kr, P(E))
    return E


def hel_coaxial(Ci, Ni, No, E, L, is_complex):
    Co = np.zeros(2 * (No + 1) ** 2)
    lib().orc_hel_coaxial(P(Co), No, P(_c(Ci)), Ni, P(E), L, int(is_complex), 0.0)
    return Co


def hel_rotate(Ci, N, H, ph, ch):
    Co = np.zeros(2 * (N + 1) ** 2)
    lib().orc_hel_rotate(P(Co), P(_c(Ci)), N, P(H), ph, ch, 0.0)
    return Co


def hel_direct(k, xs, q, xt, field=3, nrm=None, d=None, cutoff=0.0):
    """q, d complex arrays; returns (ntgt, 4) complex: potential, gradient"""
    xs = _c(xs)
    xt = _c(xt)
    f = np.zeros((len(xt), 8))
    qq = None if q is None else _c(np.stack([np.real(q), np.imag(q)], 1))
    dd = None if d is None else _c(np.stack([np.real(d), np.imag(d)], 1))
    lib().orc_hel_direct(k, len(xs), P(xs), P(qq), P(None if nrm is None else _c(nrm)), P(dd),
                         len(xt), P(xt), field, cutoff, P(f))
    return f[:, 0::2] + 1j * f[:, 1::2]


def fmm(problem, xs, xt, x0, D, depth, order_s, order_r, q=None, nrm=None, d=None,
        k=0.0, field=3, nthreads=1, dump_level=0):
    """Full FMM solve.  Returns (f[ntgt,4] or [ntgt,8], phase times, dumps)."""
    cfg = FmmCfg()
    cfg.problem = problem
    cfg.k = k
    for i in range(3):
        cfg.x0[i] = x0[i]
    cfg.D = D
    cfg.depth = depth
    for l in range(17):
        cfg.order_s[l] = int(order_s[l]) if l < len(order_s) else 0
        cfg.order_r[l] = int(order_r[l]) if l < len(order_r) else 0
    cfg.field = field
    cfg.nthreads = nthreads
    cfg.dump_level = dump_level
    dumps = None
    if dump_level:
        nb = 8 ** dump_level
        cw = 1 if problem == 1 else 2
        dumps = (np.zeros(nb * (order_s[dump_level] + 1) ** 2 * cw),
                 np.zeros(nb * (order_r[dump_level] + 1) ** 2 * cw))
        cfg.dump_mps = P(dumps[0])
        cfg.dump_mpr = P(dumps[1])
    xs = np.ascontiguousarray(xs, float)
    xt = np.ascontiguousarray(xt, float)
    # Laplace rows: potential, gradient and, with field bit 4, the second
    # derivatives (xx, yy, zz, xy, yz, zx)
    width = (10 if field & 4 else 4) if problem == 1 else 8
    f = np.zeros((len(xt), width))
    rc = lib().orc_fmm(ctypes.byref(cfg), len(xs), P(xs),
                       P(None if q is None else np.ascontiguousarray(q, float)),
                       P(None if nrm is None else np.ascontiguousarray(nrm, float)),
                       P(None if d is None else np.ascontiguousarray(d, float)),
                       len(xt), P(xt), P(f))
    if rc != 0:
        raise RuntimeError(f"oracle fmm failed: rc={rc}")
    return f, list(cfg.t_phase), dumps
