"""wbfmm_amd -- MI355X (gfx950) implementation of the wbfmm octree FMM path.

The product is the C-ABI library wbfmm_amd/libwbfmm.so (reference API in
include/wbfmm.h, kernel ABI in include/wbfmm_hip.h).  This module loads it
with ctypes and mirrors the reference's call sequence
(tools/wbfmm-laplace-tree-calc.c.in:380-460) for Python callers, tests and
the benchmark.  There is no fallback: if the library is missing or no GPU
is present, the calls raise.
"""

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libwbfmm.so")

WBFMM_PROBLEM_LAPLACE = 1
WBFMM_PROBLEM_HELMHOLTZ = 2
WBFMM_FIELD_POTENTIAL = 1
WBFMM_FIELD_GRADIENT = 2

_lib = None
_vp = ctypes.c_void_p
_u = ctypes.c_uint
_i = ctypes.c_int
_d = ctypes.c_double
_sz = ctypes.c_size_t


class WbfmmTree(ctypes.Structure):
    """Public part of wbfmm_tree_t (src/wbfmm.h:116-145)."""
    _fields_ = [
        ("problem", _i),
        ("boxes", _vp * 17),
        ("sorted", _i),
        ("maxpoints", _u), ("npoints", _u), ("ip", _vp), ("nq", _u),
        ("depth", _u), ("order_s", _u * 17), ("order_r", _u * 17),
        ("x", ctypes.c_char * 24), ("normals", _vp), ("points", _vp),
        ("mps", _vp * 17), ("mpr", _vp * 17),
        ("size", _sz), ("nstr", _sz), ("pstr", _sz),
        ("D", _d),
    ]


# name -> (restype, argtypes)
M2L_MAXN, M2L_MAXB, M2L_MAXU = 30, 96, 528


class M2LPlan(ctypes.Structure):
    """wbfmm_m2l_plan_t (include/wbfmm_hip.h): plan of the MFMA M2L."""
    _fields_ = [
        ("N", _i), ("C", _i), ("NU", _i), ("NB", _i), ("KT", _i), ("NTH", _i), ("S", _i),
        ("NR", _i),
        ("nk", _i * M2L_MAXB), ("koff", _i * M2L_MAXB), ("kbase", _i * M2L_MAXB),
        ("map1", _i * (16 * M2L_MAXB)), ("map2", _i * (16 * M2L_MAXB)),
        ("map3", _i * (16 * M2L_MAXB)),
        ("s1", _i * M2L_MAXU),
        ("th", _i * 343), ("rix", _i * 343), ("chi", _d * 343), ("r", _d * 343),
        ("rr", _d * 32),
    ]


def m2l_tables(N, w):
    """host copy of the M2L plan and tables of order N for box width w:
    (plan, A1 (NTH, KT, 64), A3 (NTH, KT, 64), A2 (NR, KT, 64),
    CS (343, N+1, 2))"""
    P = M2LPlan()
    rc = lib().wbfmm_hip_lap_m2l_tables(N, ctypes.byref(P), None, None, None)
    if rc != 0:
        raise RuntimeError(f"wbfmm_hip_lap_m2l_tables({N}) = {rc}")
    A1 = np.zeros((P.NTH, P.KT, 64))
    A3 = np.zeros((P.NTH, P.KT, 64))
    CS = np.zeros((343, N + 1, 2))
    A2 = np.zeros((P.NR, P.KT, 64))
    rc = lib().wbfmm_hip_lap_m2l_tables(N, ctypes.byref(P), ptr(A1), ptr(A3), ptr(CS))
    if rc == 0:
        rc = lib().wbfmm_hip_lap_m2l_coax(ctypes.byref(P), ctypes.c_double(w), ptr(A2))
    if rc != 0:
        raise RuntimeError(f"M2L tables of order {N}: status {rc}")
    return P, A1, A3, A2, CS


def hel_m2l_tables(N, L, E):
    """host copy of the Helmholtz M2L tiles of order N: (plan, A (NTH, KT,
    64) for both rotations, A2 (NR, 2, KT, 64) from the NR host S|R tables
    E of order L, each 2 (L+1)(L+2)(L+3)/6 doubles)"""
    P = M2LPlan()
    rc = lib().wbfmm_hip_lap_m2l_tables(N, ctypes.byref(P), None, None, None)
    if rc != 0:
        raise RuntimeError(f"wbfmm_hip_lap_m2l_tables({N}) = {rc}")
    E = np.ascontiguousarray(E, dtype=np.float64)
    nr = len(E)
    A = np.zeros((P.NTH, P.KT, 64))
    A2 = np.zeros((nr, 2, P.KT, 64))
    rc = lib().wbfmm_hip_hel_m2l_rot_tiles(ctypes.byref(P), ptr(A))
    if rc == 0:
        rc = lib().wbfmm_hip_hel_m2l_coax_tiles(ctypes.byref(P), L, nr, ptr(E), ptr(A2))
    if rc != 0:
        raise RuntimeError(f"Helmholtz M2L tiles of order {N}: status {rc}")
    return P, A, A2


_SIGS = {
    "wbfmm_tree_new": (ctypes.POINTER(WbfmmTree), [_vp, _d, _u]),
    "wbfmm_tree_add_points": (_i, [_vp, _vp, _sz, _vp, _sz, _u, _i]),
    "wbfmm_tree_refine": (_i, [_vp]),
    "wbfmm_tree_host_mirror": (_i, [_vp, _i]),
    "wbfmm_tree_laplace_coefficient_init": (_i, [_vp, _u, _u, _u]),
    "wbfmm_tree_laplace_leaf_expansions": (_i, [_vp, _vp, _i, _vp, _i, _i, _vp]),
    "wbfmm_shift_angle_table_init": (_i, []),
    "wbfmm_shift_operators_new": (_vp, [_u, _i, _vp]),
    "wbfmm_shift_operators_free": (_i, [_vp]),
    "wbfmm_laplace_coaxial_translate_init": (_i, [_i]),
    "wbfmm_laplace_upward_pass": (_i, [_vp, _vp, _u, _vp]),
    "wbfmm_laplace_downward_pass_ref": (_i, [_vp, _vp, _u, _vp, _i]),
    "wbfmm_laplace_downward_pass_avx": (_i, [_vp, _vp, _u, _vp, _i]),
    "wbfmm_laplace_box_field": (_i, [_vp, _u, _u, _vp, _i, _vp, _i, _u, _i, _vp, _vp, _i, _vp]),
    "wbfmm_laplace_field_targets": (_i, [_vp, _vp, _sz, _u, _vp, _i, _vp, _i, _u, _vp, _i]),
    "wbfmm_laplace_field_direct": (_i, [_vp, _i, _vp, _i, _i, _vp, _i, _vp, _i, _i, _u, _vp, _vp, _i]),
    "wbfmm_target_list_new": (_vp, [_vp, _u]),
    "wbfmm_target_list_add_points": (_i, [_vp, _vp, _sz, _u]),
    "wbfmm_target_list_coefficients_init": (_i, [_vp, _u]),
    "wbfmm_laplace_target_list_local_coefficients": (_i, [_vp, _u, _vp]),
    "wbfmm_target_list_local_field": (_i, [_vp, _vp, _i, _vp, _i, _vp, _i]),
    "wbfmm_target_list_free": (_i, [_vp]),
    "wbfmm_tree_free": (_i, [_vp]),
    "wbfmm_tree_coefficients_download": (_i, [_vp, _u, _i, _vp]),
    "wbfmm_tree_coefficients_zero": (_i, [_vp, _u]),
    "wbfmm_point_box": (ctypes.c_uint64, [_vp, _u, _vp]),
    "wbfmm_box_interaction_list_4": (_i, [_u, ctypes.c_uint64, _vp, _i]),
    "wbfmm_box_neighbours": (_i, [_u, ctypes.c_uint64, _vp]),
    "wbfmm_box_index": (ctypes.c_uint64, [ctypes.c_uint32] * 3),
    "wbfmm_truncation_number": (_i, [_vp, _d, _u, _d]),
    "wbfmm_hip_synchronize": (_i, []),
    "wbfmm_hip_sync": (_i, []),
    "wbfmm_hip_set_variant": (_i, [_i, _i]),
    "wbfmm_hip_get_variant": (_i, [_i]),
    "wbfmm_hip_init": (_i, []),
    "wbfmm_hip_timing_enable": (_i, [_i]),
    "wbfmm_hip_timing_reset": (_i, []),
    "wbfmm_hip_timing_query": (_i, [ctypes.c_char_p, ctypes.POINTER(_i), ctypes.POINTER(_d)]),
    "wbfmm_hip_last_error": (ctypes.c_char_p, []),
    "wbfmm_hip_malloc": (_i, [ctypes.POINTER(_vp), _sz]),
    "wbfmm_hip_free": (_i, [_vp]),
    "wbfmm_hip_memcpy": (_i, [_vp, _vp, _sz]),
    "wbfmm_hip_memset": (_i, [_vp, _i, _sz]),
    "wbfmm_hip_stage4": (_i, [_vp, _sz, _u, _vp]),
    "wbfmm_hip_lap_direct": (_i, [_vp, _vp, _vp, _vp, _u, _vp, _u, _i, _d, _vp, _i]),
    "wbfmm_hip_mfma_layout_probe": (_i, [_vp]),
    "wbfmm_hip_mfma_layout_probe_f32": (_i, [_vp]),
    "wbfmm_tree_set_precision": (_i, [_vp, _i]),
    "wbfmm_hip_lap_m2l_tables": (_i, [_i, _vp, _vp, _vp, _vp]),
    "wbfmm_hip_lap_m2l_coax": (_i, [_vp, _d, _vp]),
    "wbfmm_hip_hel_m2l_rot_tiles": (_i, [_vp, _vp]),
    "wbfmm_hip_hel_rot_tiles": (_i, [_vp, _i, _vp, _vp]),
    "wbfmm_hip_hel_m2l_coax_tiles": (_i, [_vp, _i, _i, _vp, _vp]),
    "wbfmm_tree_coefficient_init": (_i, [_vp, _u, _u, _u]),
    "wbfmm_tree_leaf_expansions": (_i, [_vp, _d, _vp, _i, _vp, _i, _i, _vp]),
    "wbfmm_shift_operators_coaxial_SR_init": (_i, [_vp, _d, _u, _u, _d, _vp]),
    "wbfmm_shift_operators_coaxial_SS_init": (_i, [_vp, _d, _u, _u, _d, _vp]),
    "wbfmm_upward_pass": (_i, [_vp, _vp, _u, _vp]),
    "wbfmm_downward_pass_ref": (_i, [_vp, _vp, _u, _vp, _i]),
    "wbfmm_downward_pass_avx": (_i, [_vp, _vp, _u, _vp, _i]),
    "wbfmm_tree_box_local_field": (_i, [_vp, _u, _u, _d, _vp, _vp, _i, _vp, _i, _vp, _i, _i, _u, _vp]),
    "wbfmm_helmholtz_field_targets": (_i, [_vp, _d, _vp, _sz, _u, _vp, _i, _vp, _i, _u, _vp, _i]),
    "wbfmm_total_field": (_i, [_d, _vp, _i, _vp, _i, _vp, _i, _vp, _i, _i, _i, _vp, _vp, _i]),
    "wbfmm_total_field_grad": (_i, [_d, _vp, _i, _vp, _i, _vp, _i, _vp, _i, _i, _i, _vp, _vp, _i]),
    "wbfmm_hip_hel_direct": (_i, [_vp, _vp, _vp, _vp, _u, _vp, _u, _d, _i, _d, _vp, _i]),
    "wbfmm_hip_hel_coaxial_table": (_i, [_i, _i, _d, _vp]),
}


class DeviceBuffer:
    """Device memory from the library's allocator (no torch needed)."""

    def __init__(self, nbytes):
        p = _vp()
        rc = lib().wbfmm_hip_malloc(ctypes.byref(p), nbytes)
        if rc != 0:
            raise MemoryError(lib().wbfmm_hip_last_error().decode())
        self.ptr = p.value
        self.nbytes = nbytes

    @classmethod
    def from_array(cls, a):
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        b.upload(a)
        return b

    def upload(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        if lib().wbfmm_hip_memcpy(self.ptr, a.ctypes.data, a.nbytes) != 0:
            raise RuntimeError(lib().wbfmm_hip_last_error().decode())

    def download(self, a):
        assert a.flags["C_CONTIGUOUS"] and a.nbytes <= self.nbytes
        if lib().wbfmm_hip_memcpy(a.ctypes.data, self.ptr, a.nbytes) != 0:
            raise RuntimeError(lib().wbfmm_hip_last_error().decode())
        return a

    def zero(self):
        lib().wbfmm_hip_memset(self.ptr, 0, self.nbytes)

    def free(self):
        if self.ptr:
            lib().wbfmm_hip_free(self.ptr)
            self.ptr = None


def lib():
    """Load libwbfmm.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} not built: run python -m wbfmm_amd.build")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def ptr(a):
    """host numpy array, torch tensor or int address -> void pointer."""
    if a is None:
        return None
    if isinstance(a, int):
        return a
    if isinstance(a, DeviceBuffer):
        return a.ptr
    if isinstance(a, np.ndarray):
        assert a.flags["C_CONTIGUOUS"]
        return a.ctypes.data
    return a.data_ptr()  # torch tensor (device memory)


def sync():
    rc = lib().wbfmm_hip_synchronize()
    if rc != 0:
        raise RuntimeError(lib().wbfmm_hip_last_error().decode())


class Timing:
    """Per-kernel HIP-event timing on the library stream."""

    @staticmethod
    def enable(on=True):
        lib().wbfmm_hip_timing_enable(1 if on else 0)

    @staticmethod
    def reset():
        lib().wbfmm_hip_timing_reset()

    @staticmethod
    def query(name):
        n = _i()
        ms = _d()
        lib().wbfmm_hip_timing_query(name.encode(), ctypes.byref(n), ctypes.byref(ms))
        return n.value, ms.value


class LaplaceSolver:
    """Laplace FMM through the drop-in API, in the reference driver's call
    order (tools/wbfmm-laplace-tree-calc.c.in:380-460).

    xs: source positions (n,3) float64, host numpy or device torch tensor
    (then pass its row stride in bytes via pstr).  Targets given at
    construction are registered as a target list so the downward pass runs
    over boxes holding sources or targets only.
    """

    def __init__(self, xs, x0, D, depth, order_s, order_r, targets=None,
                 pstr=24, tstr=24, ntgt=None, host_mirror=True, nsrc=None,
                 normals=None, nstr=24, precision=64, nq=1):
        L = lib()
        if precision == 32:
            host_mirror = False
        self.nq = nq   # source components (vector sources: nq = 3)
        self.depth = depth
        self.nsrc = len(xs) if nsrc is None else nsrc
        x0 = np.ascontiguousarray(x0, dtype=np.float64)
        self._x0 = x0
        self.xs = xs
        self.tree = L.wbfmm_tree_new(ptr(x0), float(D), self.nsrc)
        if not self.tree:
            raise RuntimeError("wbfmm_tree_new failed")
        self.t = ctypes.cast(self.tree, _vp).value
        L.wbfmm_tree_host_mirror(self.t, 1 if host_mirror else 0)
        if precision != 64:
            L.wbfmm_tree_set_precision(self.t, precision)
        lmax = max(max(order_s[2:depth + 1]), max(order_r[2:depth + 1]))
        L.wbfmm_shift_angle_table_init()
        self.op = L.wbfmm_shift_operators_new(lmax, 0, None)
        L.wbfmm_laplace_coaxial_translate_init(lmax + 2)
        self.order_s = list(order_s)
        self.order_r = list(order_r)
        self.targets = targets
        self.tstr = tstr
        self.ntgt = ntgt if ntgt is not None else (None if targets is None else len(targets))
        self.pstr = pstr
        self.normals = normals
        self.nstr = nstr
        self.tl = None

    def build_tree(self):
        """add_points + refine x depth + coefficient init (+ target list)."""
        L = lib()
        L.wbfmm_tree_add_points(self.t, ptr(self.xs), self.pstr, ptr(self.normals),
                                self.nstr if self.normals is not None else 0, self.nsrc, 0)
        for _ in range(self.depth):
            L.wbfmm_tree_refine(self.t)
        self.tree.contents.problem = WBFMM_PROBLEM_LAPLACE
        self.tree.contents.nq = self.nq
        for l in range(1, self.depth + 1):
            L.wbfmm_tree_laplace_coefficient_init(self.t, l, self.order_r[l], self.order_s[l])
        if self.targets is not None:
            if self.tl is None:
                self.tl = L.wbfmm_target_list_new(self.t, self.ntgt)
            L.wbfmm_target_list_add_points(self.tl, ptr(self.targets), self.tstr, self.ntgt)
        self._fresh = True   # coefficient_init leaves every level zeroed

    def passes(self, q, sstr=1, d=None, dstr=1):
        """P2M, upward pass depth..3, downward pass 2..depth (the levels
        are zeroed first unless build_tree has just initialised them)."""
        L = lib()
        if not getattr(self, "_fresh", False):
            for l in range(1, self.depth + 1):
                L.wbfmm_tree_coefficients_zero(self.t, l)
        self._fresh = False
        L.wbfmm_tree_laplace_leaf_expansions(self.t, ptr(q), sstr, ptr(d), dstr, 1, None)
        for l in range(self.depth, 2, -1):
            L.wbfmm_laplace_upward_pass(self.t, self.op, l, None)
        for l in range(2, self.depth + 1):
            L.wbfmm_laplace_downward_pass_ref(self.t, self.op, l, None, 0)

    def field(self, q, f, field=3, fstr=4, sstr=1, d=None, dstr=1):
        """target-list field (L2P + near field) into f (accumulated)."""
        L = lib()
        L.wbfmm_target_list_coefficients_init(self.tl, field)
        L.wbfmm_laplace_target_list_local_coefficients(self.tl, 1, None)
        L.wbfmm_target_list_local_field(self.tl, ptr(q), sstr, ptr(d), dstr, ptr(f), fstr)

    def field_points(self, xt, q, f, field=3, fstr=4, sstr=1, d=None, dstr=1, tstr=24,
                     ntgt=None):
        """field at arbitrary points (wbfmm_laplace_field_targets): with nq
        components, field 1/2/3 gives the potentials and gradients of every
        component, 4 the curl, 6 the curl and its gradient"""
        n = len(xt) if ntgt is None else ntgt
        rc = lib().wbfmm_laplace_field_targets(self.t, ptr(xt), tstr, n, ptr(q), sstr, ptr(d),
                                               dstr, field, ptr(f), fstr)
        if rc != 0:
            raise RuntimeError(f"wbfmm_laplace_field_targets = {rc}")

    def solve(self, q, f, field=3, fstr=4, d=None):
        self.build_tree()
        self.passes(q, d=d)
        self.field(q, f, field, fstr, d=d)

    def coefficients(self, level, which):
        """download mps (which=0) or mpr (which=1) of a level as an
        (8^level, C) array in box-major order."""
        C = (self.order_s[level] + 1) ** 2 if which == 0 else (self.order_r[level] + 1) ** 2
        nb = 8 ** level
        nq = self.nq
        buf = np.zeros(nb * C * nq)
        rc = lib().wbfmm_tree_coefficients_download(self.t, level, which, ptr(buf))
        if rc != 0:
            raise RuntimeError("download failed")
        # [group of 8][coef][child][component] -> [box][coef] ([component][box][coef])
        a = buf.reshape(nb // 8, C, 8, nq).transpose(3, 0, 2, 1).reshape(nq, nb, C)
        return a[0] if nq == 1 else a

    def close(self):
        L = lib()
        if self.tl:
            L.wbfmm_target_list_free(self.tl)
            self.tl = None
        if self.t:
            L.wbfmm_tree_free(self.t)
            self.t = None
        if self.op:
            L.wbfmm_shift_operators_free(self.op)
            self.op = None


class HelmholtzSolver:
    """Helmholtz FMM through the drop-in API, in the reference driver's
    call order (tools/wbfmm-tree-calc.c.in:306-417): shift operators,
    S|R tables for levels 2..depth and S|S tables for levels 3..depth,
    tree, coefficients, P2M, upward pass, downward pass, field at the
    targets.  Strengths are complex: (n, 2) float64 arrays (re, im)."""

    def __init__(self, xs, x0, D, depth, order_s, order_r, k, pstr=24,
                 normals=None, nstr=24, nsrc=None, targets=None, ntgt=None, tstr=24,
                 precision=64):
        L = lib()
        self.targets = targets
        self.ntgt = ntgt if ntgt is not None else (None if targets is None else len(targets))
        self.tstr = tstr
        self.tl = None
        self.depth = depth
        self.k = float(k)
        self.nsrc = len(xs) if nsrc is None else nsrc
        self._x0 = np.ascontiguousarray(x0, dtype=np.float64)
        self.xs = xs
        self.pstr = pstr
        self.normals = normals
        self.nstr = nstr
        self.order_s = list(order_s)
        self.order_r = list(order_r)
        self.tree = L.wbfmm_tree_new(ptr(self._x0), float(D), self.nsrc)
        if not self.tree:
            raise RuntimeError("wbfmm_tree_new failed")
        self.t = ctypes.cast(self.tree, _vp).value
        L.wbfmm_tree_host_mirror(self.t, 0)
        if precision != 64:
            L.wbfmm_tree_set_precision(self.t, precision)
        lmax = max(max(order_s[2:depth + 1]), max(order_r[2:depth + 1]))
        L.wbfmm_shift_angle_table_init()
        self.op = L.wbfmm_shift_operators_new(lmax, 0, None)
        for l in range(2, depth + 1):
            lo = max(order_s[l], order_r[l])
            if L.wbfmm_shift_operators_coaxial_SR_init(self.op, float(D), l, lo, self.k, None):
                raise RuntimeError("S|R init failed")
        for l in range(3, depth + 1):
            lo = max(order_s[l], order_s[l - 1], order_r[l], order_r[l - 1])
            if L.wbfmm_shift_operators_coaxial_SS_init(self.op, float(D), l, lo, self.k, None):
                raise RuntimeError("S|S init failed")

    def build_tree(self):
        L = lib()
        L.wbfmm_tree_add_points(self.t, ptr(self.xs), self.pstr, ptr(self.normals),
                                self.nstr if self.normals is not None else 0, self.nsrc, 0)
        for _ in range(self.depth):
            L.wbfmm_tree_refine(self.t)
        self.tree.contents.problem = WBFMM_PROBLEM_HELMHOLTZ
        self.tree.contents.nq = 1
        for l in range(1, self.depth + 1):
            L.wbfmm_tree_coefficient_init(self.t, l, self.order_r[l], self.order_s[l])
        if self.targets is not None:
            # registered targets restrict the downward pass to boxes holding
            # sources or targets
            if self.tl is None:
                self.tl = L.wbfmm_target_list_new(self.t, self.ntgt)
            L.wbfmm_target_list_add_points(self.tl, ptr(self.targets), self.tstr, self.ntgt)
        self._fresh = True   # coefficient_init leaves every level zeroed

    def passes(self, q, sstr=2, d=None, dstr=2):
        L = lib()
        if not getattr(self, "_fresh", False):
            for l in range(1, self.depth + 1):
                L.wbfmm_tree_coefficients_zero(self.t, l)
        self._fresh = False
        L.wbfmm_tree_leaf_expansions(self.t, self.k, ptr(q), sstr, ptr(d), dstr, 1, None)
        for l in range(self.depth, 2, -1):
            L.wbfmm_upward_pass(self.t, self.op, l, None)
        for l in range(2, self.depth + 1):
            L.wbfmm_downward_pass_ref(self.t, self.op, l, None, 0)

    def field(self, xt, q, f, field=3, fstr=8, tstr=24, ntgt=None, sstr=2, d=None, dstr=2):
        n = len(xt) if ntgt is None else ntgt
        rc = lib().wbfmm_helmholtz_field_targets(self.t, self.k, ptr(xt), tstr, n, ptr(q), sstr,
                                                 ptr(d), dstr, field, ptr(f), fstr)
        if rc != 0:
            raise RuntimeError(f"wbfmm_helmholtz_field_targets = {rc}")

    def coefficients(self, level, which):
        """mps (0) or mpr (1) of a level as an (8^level, C) complex array"""
        N = self.order_s[level] if which == 0 else self.order_r[level]
        C = (N + 1) ** 2
        nb = 8 ** level
        buf = np.zeros(nb * C * 2)
        if lib().wbfmm_tree_coefficients_download(self.t, level, which, ptr(buf)) != 0:
            raise RuntimeError("download failed")
        # [group of 8][coef][child][re, im] -> [box][coef]
        z = buf[0::2] + 1j * buf[1::2]
        return z.reshape(nb // 8, C, 8).transpose(0, 2, 1).reshape(nb, C)

    def close(self):
        L = lib()
        if self.tl:
            L.wbfmm_target_list_free(self.tl)
            self.tl = None
        if self.t:
            L.wbfmm_tree_free(self.t)
            self.t = None
        if self.op:
            L.wbfmm_shift_operators_free(self.op)
            self.op = None
