"""The C-ABI library loads and exports every function the public headers
declare (no GPU needed: nothing is called)."""

import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"#define[^\n]*(\\\n[^\n]*)*", "", text)
    names = set(re.findall(r"\b((?:wbfmm)_[A-Za-z0-9_]+)\s*\(", text))
    return {n for n in names if not n.startswith("wbfmm_element_") and
            not n.endswith("_t")}


def test_library_exports_declared_symbols(wb):
    lib = ctypes.CDLL(wb.LIB_PATH)
    missing = []
    for h in ("wbfmm.h", "wbfmm_hip.h", "wbfmm_serve.h"):
        for name in sorted(declared(h)):
            try:
                getattr(lib, name)
            except AttributeError:
                missing.append((h, name))
    assert not missing, missing


def test_headers_compile_as_c(tmp_path):
    import subprocess
    src = tmp_path / "t.c"
    src.write_text('#include "wbfmm.h"\n#include "wbfmm_hip.h"\n#include "wbfmm_serve.h"\n'
                   'int main(void){return 0;}\n')
    r = subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                        "-c", str(src), "-o", str(tmp_path / "t.o")], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_kernel_variant_switch(wb):
    # host-side state of the calling thread only (no device work): the
    # M2L variant takes -1..2, the field variant -1..1, anything else is
    # refused and leaves the setting as it was
    lib = ctypes.CDLL(wb.LIB_PATH)
    try:
        for which, vmax in ((0, 2), (1, 1)):
            for v in range(-1, vmax + 1):
                assert lib.wbfmm_hip_set_variant(which, v) == 0
                assert lib.wbfmm_hip_get_variant(which) == v
            assert lib.wbfmm_hip_set_variant(which, vmax + 1) != 0
            assert lib.wbfmm_hip_set_variant(which, -2) != 0
            assert lib.wbfmm_hip_get_variant(which) == vmax
        assert lib.wbfmm_hip_set_variant(2, 0) != 0
        assert lib.wbfmm_hip_get_variant(2) == -1
    finally:
        lib.wbfmm_hip_set_variant(0, -1)
        lib.wbfmm_hip_set_variant(1, -1)
