"""Python view of the serving engine (include/wbfmm_serve.h).

Server queues independent FMM solves (Laplace or Helmholtz) that the
library runs side by side on a pool of HIP streams; results come back per
ticket with their timings and, when asked, the relative error of a sample
of targets against the FP64 direct sum.
"""

import ctypes

import numpy as np

from . import (WBFMM_FIELD_GRADIENT, WBFMM_FIELD_POTENTIAL, WBFMM_PROBLEM_HELMHOLTZ,
               WBFMM_PROBLEM_LAPLACE, lib, ptr)

_vp = ctypes.c_void_p
_i = ctypes.c_int
_u = ctypes.c_uint
_d = ctypes.c_double
_sz = ctypes.c_size_t
_i64 = ctypes.c_int64
_u64 = ctypes.c_uint64


class Request(ctypes.Structure):
    _fields_ = [
        ("problem", _i), ("precision", _i), ("k", _d), ("tol", _d),
        ("nsrc", _u), ("xs", _vp), ("pstr", _sz), ("q", _vp),
        ("normals", _vp), ("nstr", _sz), ("dip", _vp),
        ("ntgt", _u), ("xt", _vp), ("tstr", _sz),
        ("field", _u), ("f", _vp), ("fstr", _i),
        ("depth", _i), ("order", _i), ("check", _u),
    ]


class Result(ctypes.Structure):
    _fields_ = [
        ("status", _i), ("depth", _i), ("order", _i), ("precision", _i),
        ("queue_ms", _d), ("tables_ms", _d), ("solve_ms", _d),
        ("latency_ms", _d), ("check_ms", _d), ("footprint", _sz),
        ("err_potential", _d), ("err_gradient", _d),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class Stats(ctypes.Structure):
    _fields_ = [
        ("completed", _u64), ("targets", _u64),
        ("p50_ms", _d), ("p99_ms", _d), ("mean_ms", _d), ("max_ms", _d),
        ("span_s", _d), ("peak_bytes", _sz),
    ]


_SIGS = {
    "wbfmm_serve_new": (_vp, [_i, _sz]),
    "wbfmm_serve_submit": (_i64, [_vp, ctypes.POINTER(Request)]),
    "wbfmm_serve_wait": (_i, [_vp, _i64, ctypes.POINTER(Result)]),
    "wbfmm_serve_stats": (_i, [_vp, ctypes.POINTER(Stats)]),
    "wbfmm_serve_reset_stats": (_i, [_vp]),
    "wbfmm_serve_plan": (_i, [_vp, ctypes.POINTER(Request), ctypes.POINTER(_i),
                              ctypes.POINTER(_i), ctypes.POINTER(_sz)]),
    "wbfmm_serve_tune": (_i, [_vp, ctypes.POINTER(Request), ctypes.POINTER(_i), _i,
                              ctypes.POINTER(_i)]),
    "wbfmm_serve_tuned": (_i, [_vp, ctypes.POINTER(Request), ctypes.POINTER(_i),
                               ctypes.POINTER(_i), ctypes.POINTER(_i)]),
    "wbfmm_serve_free": (_i, [_vp]),
    "wbfmm_serve_debug_fail_allocations": (_i, [_vp, _i]),
}
_bound = False


def _L():
    global _bound
    L = lib()
    if not _bound:
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _bound = True
    return L


def field_width(problem, field):
    c = 2 if problem == WBFMM_PROBLEM_HELMHOLTZ else 1
    return c * ((1 if field & WBFMM_FIELD_POTENTIAL else 0) + (3 if field & WBFMM_FIELD_GRADIENT else 0))


def make_request(problem, xs, q=None, xt=None, k=0.0, tol=1e-6, field=3, f=None, fstr=None,
                 normals=None, dip=None, depth=0, order=0, check=0, nsrc=None, ntgt=None,
                 pstr=24, tstr=24, nstr=24, precision=64):
    """Request over numpy arrays (host) or torch tensors (device); f is
    the output array (accumulated).  Keep the arrays alive until wait."""
    r = Request()
    r.problem = problem
    r.precision = precision
    r.k = k
    r.tol = tol
    r.nsrc = len(xs) if nsrc is None else nsrc
    r.xs = ptr(xs)
    r.pstr = pstr
    r.q = ptr(q)
    r.normals = ptr(normals)
    r.nstr = nstr
    r.dip = ptr(dip)
    r.ntgt = 0 if xt is None else (len(xt) if ntgt is None else ntgt)
    r.xt = ptr(xt)
    r.tstr = tstr
    r.field = field
    r.f = ptr(f)
    r.fstr = field_width(problem, field) if fstr is None else fstr
    r.depth = depth
    r.order = order
    r.check = check
    return r


class Server:
    def __init__(self, nstreams=4, budget=0):
        self.s = _L().wbfmm_serve_new(nstreams, budget)
        if not self.s:
            raise RuntimeError("wbfmm_serve_new failed: " + lib().wbfmm_hip_last_error().decode())
        self._keep = {}

    def submit(self, req, keep=None):
        t = _L().wbfmm_serve_submit(self.s, ctypes.byref(req))
        if t < 0:
            raise RuntimeError(f"wbfmm_serve_submit = {t}")
        self._keep[t] = (req, keep)
        return t

    def wait(self, ticket):
        res = Result()
        _L().wbfmm_serve_wait(self.s, ticket, ctypes.byref(res))
        self._keep.pop(ticket, None)
        return res

    def plan(self, req):
        d, o, fp = _i(), _i(), _sz()
        rc = _L().wbfmm_serve_plan(self.s, ctypes.byref(req), ctypes.byref(d), ctypes.byref(o),
                                   ctypes.byref(fp))
        if rc:
            raise RuntimeError(f"wbfmm_serve_plan = {rc}")
        return d.value, o.value, fp.value

    def tune(self, req, depths):
        arr = (_i * len(depths))(*depths)
        best = _i()
        rc = _L().wbfmm_serve_tune(self.s, ctypes.byref(req), arr, len(depths), ctypes.byref(best))
        if rc:
            raise RuntimeError(f"wbfmm_serve_tune = {rc}")
        return best.value

    def tuned(self, req):
        """the autotuner's choice for req's class: (depth, M2L variant,
        field variant); depth 0 and variants -1 when the class is not
        tuned"""
        d, vm, vf = _i(), _i(), _i()
        rc = _L().wbfmm_serve_tuned(self.s, ctypes.byref(req), ctypes.byref(d), ctypes.byref(vm),
                                    ctypes.byref(vf))
        if rc:
            raise RuntimeError(f"wbfmm_serve_tuned = {rc}")
        return d.value, vm.value, vf.value

    def debug_fail_allocations(self, count):
        """the next submitted request has its first `count` device
        allocations refused (failure-handling tests)"""
        rc = _L().wbfmm_serve_debug_fail_allocations(self.s, count)
        if rc:
            raise RuntimeError(f"wbfmm_serve_debug_fail_allocations = {rc}")

    def stats(self):
        st = Stats()
        _L().wbfmm_serve_stats(self.s, ctypes.byref(st))
        return st

    def reset_stats(self):
        _L().wbfmm_serve_reset_stats(self.s)

    def close(self):
        if self.s:
            _L().wbfmm_serve_free(self.s)
            self.s = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


__all__ = ["Server", "Request", "Result", "Stats", "make_request", "field_width",
           "WBFMM_PROBLEM_LAPLACE", "WBFMM_PROBLEM_HELMHOLTZ"]
