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

HIP_SOURCES = ["runtime.hip", "laplace.hip", "m2l_mfma.hip", "hel_m2l_mfma.hip", "helmholtz.hip"]
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


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    jobs = []
    for s in HIP_SOURCES:
        o = os.path.join(OBJ, s + ".o")
        jobs.append([HIPCC, *HIP_FLAGS, "-c", "-o", o, os.path.join(CSRC, s)])
    for s in C_SOURCES:
        o = os.path.join(OBJ, s + ".o")
        jobs.append([CC, *C_FLAGS, "-c", "-o", o, os.path.join(CSRC, s)])
    for s in CXX_SOURCES:
        o = os.path.join(OBJ, s + ".o")
        jobs.append([CXX, *CXX_FLAGS, "-c", "-o", o, os.path.join(CSRC, s)])
    with ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        for r in ex.map(_run, jobs):
            if verbose and (r.stdout or r.stderr):
                sys.stderr.write(r.stdout + r.stderr)
    objs = [os.path.join(OBJ, s + ".o") for s in HIP_SOURCES + C_SOURCES + CXX_SOURCES]
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", "-o", tmp, *objs,
          "-lquadmath", "-lm", "-lpthread"])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
