"""Build libwbfmm.so (gfx950) in-tree.

hipcc compiles the HIP kernels and launch layer, gcc the C host layer
(the drop-in API and the FP64 table builders), g++ the serving engine;
hipcc links the shared library into wbfmm_amd/libwbfmm.so.  Run as `python -m wbfmm_amd.build`.
"""

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INC = os.path.join(ROOT, "include")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libwbfmm.so")

# the longest compiles first (the build runs 8 at a time)
HIP_SOURCES = [
    # the Helmholtz and Laplace matrix-core kernels, one range of orders per unit
    "hel_m2l_o26_27.hip", "hel_m2l_o24_25.hip", "hel_m2l_o22_23.hip", "hel_m2l_o19_21.hip",
    "m2l_lap_o22_24.hip", "m2l_lap_o19_21.hip", "m2l_lap_o16_18.hip", "hel_m2l_o15_18.hip",
    "m2l_lap_o13_15.hip", "m2l_lap_o11_12.hip", "m2l_lap_o08_10.hip", "hel_m2l_o10_14.hip",
    "m2l_lap_o02_07.hip", "hel_m2l_o02_09.hip",
    "helmholtz.hip", "laplace.hip", "m2l_mfma.hip", "hel_m2l_mfma.hip", "runtime.hip",
]
C_SOURCES = ["wbfmm_api.c", "wbfmm_tables.c"]
CXX_SOURCES = ["serve.cpp"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CC = os.environ.get("CC", "gcc")
CXX = os.environ.get("CXX", "g++")

HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950",
    # keep the compiler from emitting packed FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-Wno-unused-value", "-I", INC, "-I", CSRC,
]
C_FLAGS = ["-O2", "-g", "-fPIC", "-std=gnu11", "-Wall", "-Wno-unused-function",
           "-I", INC]
CXX_FLAGS = ["-O2", "-g", "-fPIC", "-std=c++17", "-Wall", "-pthread", "-I", INC]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {os.path.basename(cmd[-1])}")
    return r


def _deps(obj):
    """files the object was built from, per the compiler's -MMD list"""
    dep = obj + ".d"
    if not os.path.exists(dep):
        return None
    with open(dep) as fh:
        text = fh.read().replace("\\\n", " ")
    if ":" not in text:
        return None
    return [f for f in text.split(":", 1)[1].split() if f]


def _stale(src, obj):
    """object missing or older than a file it was built from"""
    if not os.path.exists(obj):
        return True
    t = os.path.getmtime(obj)
    deps = _deps(obj) or [src]
    return any(not os.path.exists(f) or os.path.getmtime(f) > t for f in deps + [src])


def build(verbose=False, incremental=False):
    """compile every source (incremental: only those whose object is out
    of date) and link the library"""
    os.makedirs(OBJ, exist_ok=True)
    jobs = []
    for srcs, cmd in ((HIP_SOURCES, [HIPCC, *HIP_FLAGS]), (C_SOURCES, [CC, *C_FLAGS]),
                      (CXX_SOURCES, [CXX, *CXX_FLAGS])):
        for s in srcs:
            o = os.path.join(OBJ, s + ".o")
            src = os.path.join(CSRC, s)
            if incremental and not _stale(src, o):
                continue
            jobs.append([*cmd, "-MMD", "-MF", o + ".d", "-c", "-o", o, src])
    with ThreadPoolExecutor(max_workers=max(1, min(8, len(jobs)))) as ex:
        for r in ex.map(_run, jobs):
            if verbose and (r.stdout or r.stderr):
                sys.stderr.write(r.stdout + r.stderr)
    objs = [os.path.join(OBJ, s + ".o") for s in HIP_SOURCES + C_SOURCES + CXX_SOURCES]
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", "-o", tmp, *objs,
          "-lquadmath", "-lm", "-lpthread"])
    os.replace(tmp, LIB)
    build_drivers()
    return LIB


DRIVERS = ["wbfmm-tree-calc", "wbfmm-laplace-tree-calc"]
BIN = os.path.join(HERE, "bin")


def build_drivers():
    """the replacement driver programs, plain C against include/wbfmm.h,
    linked to the library (found next to bin/ at run time)"""
    os.makedirs(BIN, exist_ok=True)
    src_dir = os.path.join(HERE, "drivers")
    for d in DRIVERS:
        _run([CC, "-O2", "-g", "-std=gnu11", "-Wall", "-I", INC, "-I", src_dir, "-o",
              os.path.join(BIN, d), os.path.join(src_dir, d + ".c"), "-L", HERE, "-l:libwbfmm.so",
              "-Wl,-rpath,$ORIGIN/..", "-lm"])


if __name__ == "__main__":
    print(build(verbose=True, incremental="--incremental" in sys.argv))
