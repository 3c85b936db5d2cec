"""The C-ABI library loads and exports every function declared in include/*.h
(no compute calls: this runs without a GPU)."""
import ctypes
import glob
import os
import re

from sqdft_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        src = open(h).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        for m in re.finditer(r"^\s*(?:int|void|double|long)\s+\**\s*(\w+)\s*\(", src, flags=re.M):
            names.add(m.group(1))
    return sorted(names)


def test_headers_declare_the_api():
    names = declared_functions()
    for n in ("sqg_create", "sqg_density", "sqg_forces_nonlocal", "sqg_hsub_apply_batch",
              "sqdft_system_create"):
        assert n in names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_bindings_cover_the_headers():
    assert set(declared_functions()) <= set(_lib.SIGNATURES)
