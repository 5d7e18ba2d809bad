"""The C ABI library loads and exports every symbol include/*.h declares
(no compute call: this runs without a GPU)."""
import ctypes
import os
import subprocess

import pytest

N = pytest.importorskip("chuanflow_amd.native")


def test_every_declared_symbol_is_exported():
    names = N.declared_symbols()
    assert len(names) > 40
    lib = ctypes.CDLL(N.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_headers_compile_as_c():
    inc = N.INCLUDE_DIR
    src = "#include <cfx.h>\n#include <cfx_host.h>\nint main(void){return 0;}\n"
    r = subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-I" + inc, "-x", "c", "-", "-o", os.devnull],
                       input=src, text=True, capture_output=True)
    assert r.returncode == 0, r.stderr


def test_library_has_gfx950_code_object():
    data = open(N.LIB_PATH, "rb").read()
    assert b"gfx950" in data
