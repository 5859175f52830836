"""The allocation-free M2L entry point (wbfmm_hip_lap_m2l_mfma_noalloc)
against the allocating one (wbfmm_hip_lap_m2l_mfma) on the same level
data, through the C ABI.

Three calls in a row, each smaller than the one before: a level-6 set of
targets that runs the packed sparse-level kernel, a level-4 set that runs
the split form (partials in the library's scratch), and a level-3 set at
a lower order.  Both entry points run the same kernels in the same order,
so the results agree to rounding (1e-13 of the largest coefficient is the
bar written below).  Sizes beyond the scratch are refused before any
launch, with a message.
"""

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

_vp, _u, _i, _d = ctypes.c_void_p, ctypes.c_uint, ctypes.c_int, ctypes.c_double
U8 = ctypes.c_uint * 8
M2L_ARGS = [_vp, ctypes.POINTER(_u), ctypes.POINTER(_u), _vp, _vp, _u, _d, _i, _vp, _vp,
            _vp, _vp, _vp, _vp, _vp, _vp]


def _bind(L):
    L.wbfmm_hip_split_classes.restype = _i
    L.wbfmm_hip_split_classes.argtypes = [_vp, _u, _vp, ctypes.POINTER(_u), ctypes.POINTER(_u)]
    L.wbfmm_hip_lap_m2l_plan.restype = _i
    L.wbfmm_hip_lap_m2l_plan.argtypes = [_i, _vp, ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                                        ctypes.POINTER(_vp), ctypes.POINTER(_vp)]
    for name in ("wbfmm_hip_lap_m2l_mfma", "wbfmm_hip_lap_m2l_mfma_noalloc"):
        fn = getattr(L, name)
        fn.restype = _i
        fn.argtypes = M2L_ARGS


class Level:
    """device copy of one level: point ranges of every box, random source
    expansions, and the occupied boxes split by Morton child digit"""

    def __init__(self, wb, rng, level, N, fill):
        nb = 8 ** level
        self.level, self.N, self.C = level, N, (N + 1) ** 2
        occ = rng.random(nb) < fill
        counts = np.where(occ, rng.integers(1, 4, nb), 0).astype(np.uint32)
        end = np.cumsum(counts).astype(np.uint32)
        start = (end - counts).astype(np.uint32)
        mps = rng.normal(size=(nb, self.C))
        mps[~occ] = 0.0
        tl = np.nonzero(occ)[0].astype(np.uint32)
        self.ntgt = len(tl)
        self.nb = nb
        self.d_start = wb.DeviceBuffer.from_array(start)
        self.d_end = wb.DeviceBuffer.from_array(end)
        self.d_mps = wb.DeviceBuffer.from_array(mps)
        d_tl = wb.DeviceBuffer.from_array(tl)
        self.d_cls = wb.DeviceBuffer(4 * max(self.ntgt, 1))
        self.cstart, self.ccount = U8(), U8()
        L = wb.lib()
        assert L.wbfmm_hip_split_classes(d_tl.ptr, self.ntgt, self.d_cls.ptr, self.cstart,
                                         self.ccount) == 0
        wb.sync()
        d_tl.free()
        self.nchunks = sum((self.ccount[c] + 15) // 16 for c in range(8))

    def free(self):
        for b in (self.d_start, self.d_end, self.d_mps, self.d_cls):
            b.free()


class Tables:
    def __init__(self, wb, N, w):
        L = wb.lib()
        self.P = wb.M2LPlan()
        self.plan_dev, self.A1, self.A3, self.CS = _vp(), _vp(), _vp(), _vp()
        assert L.wbfmm_hip_lap_m2l_plan(N, ctypes.byref(self.P), ctypes.byref(self.plan_dev),
                                        ctypes.byref(self.A1), ctypes.byref(self.A3),
                                        ctypes.byref(self.CS)) == 0
        A2 = np.zeros((self.P.NR, self.P.KT, 64))
        assert L.wbfmm_hip_lap_m2l_coax(ctypes.byref(self.P), ctypes.c_double(w), A2.ctypes.data) == 0
        self.d_A2 = wb.DeviceBuffer.from_array(A2)


def run(wb, name, lv, T, D):
    L = wb.lib()
    out = wb.DeviceBuffer(8 * lv.nb * lv.C)
    out.zero()
    rc = getattr(L, name)(lv.d_cls.ptr, lv.cstart, lv.ccount, lv.d_start.ptr, lv.d_end.ptr,
                          lv.level, D, lv.N, lv.d_mps.ptr, out.ptr, T.plan_dev, T.A1, T.A3,
                          T.d_A2.ptr, T.CS, ctypes.byref(T.P))
    assert rc == 0, (name, rc, L.wbfmm_hip_last_error())
    wb.sync()
    host = out.download(np.zeros((lv.nb, lv.C)))
    out.free()
    return host


@pytest.fixture(scope="module")
def lib(wb):
    _bind(wb.lib())
    return wb


# (level, order, fill): each call smaller than the one before
CASES = [(6, 12, 0.35), (4, 12, 0.6), (3, 8, 0.7)]


def test_noalloc_m2l_matches_allocating_entry(lib):
    wb = lib
    rng = np.random.default_rng(2024)
    D = 1.0
    for level, N, fill in CASES:
        print(f"no-alloc M2L: level {level} order {N} fill {fill}", flush=True)
        lv = Level(wb, rng, level, N, fill)
        T = Tables(wb, N, D / 2 ** level)
        ref = run(wb, "wbfmm_hip_lap_m2l_mfma", lv, T, D)
        got = run(wb, "wbfmm_hip_lap_m2l_mfma_noalloc", lv, T, D)
        scale = np.abs(ref).max()
        assert scale > 0
        err = np.abs(got - ref).max() / scale
        print(f"  targets {lv.ntgt} chunks {lv.nchunks} max rel diff {err:.2e}", flush=True)
        assert err <= 1e-13
        lv.free()


def test_noalloc_m2l_refuses_sizes_beyond_scratch(lib):
    """refused on the host before any launch: a level above 8 (the
    occupancy bitmap region), and split partials above the scratch"""
    wb = lib
    L = wb.lib()
    dummy = wb.DeviceBuffer(256)
    T = Tables(wb, 16, 1.0 / 64)
    cstart, ccount = U8(), U8()
    ccount[0] = 16
    rc = L.wbfmm_hip_lap_m2l_mfma_noalloc(dummy.ptr, cstart, ccount, dummy.ptr, dummy.ptr, 9,
                                          1.0, 16, dummy.ptr, dummy.ptr, T.plan_dev, T.A1, T.A3,
                                          T.d_A2.ptr, T.CS, ctypes.byref(T.P))
    assert rc != 0 and b"level 9" in L.wbfmm_hip_last_error()
    # 1000 chunks of 16 targets at order 16: two blocks per chunk, so
    # 2 x 16000 x 289 x 8 bytes = 74 MB of partials
    ccount[0] = 16000
    rc = L.wbfmm_hip_lap_m2l_mfma_noalloc(dummy.ptr, cstart, ccount, dummy.ptr, dummy.ptr, 6,
                                          1.0, 16, dummy.ptr, dummy.ptr, T.plan_dev, T.A1, T.A3,
                                          T.d_A2.ptr, T.CS, ctypes.byref(T.P))
    assert rc != 0 and b"scratch region" in L.wbfmm_hip_last_error()
    dummy.free()
