"""The C-ABI library loads without a GPU and exports every function that
include/*.h declares (no compute calls)."""
import ctypes
import glob
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DECL = re.compile(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b([A-Za-z_]\w*)\s*\(", re.M)


def declared(path):
    text = open(path).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    text = re.sub(r"typedef\s+struct\s*\{.*?\}\s*\w+\s*;", "", text, flags=re.S)
    names = set()
    for m in DECL.finditer(text):
        name = m.group(1)
        if name in ("if", "while", "for", "return", "sizeof", "defined"):
            continue
        names.add(name)
    return names


def test_headers_declare_the_entry_points():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        names |= declared(h)
    assert "atsim_sim_frame" in names and "atsim_fresnel" in names and "atsim_fft2d" in names


def test_library_exports_every_declared_symbol():
    from atmosphericturbulencesimul_amd._lib import LIB_PATH
    if not os.path.exists(LIB_PATH):
        pytest.fail(f"{LIB_PATH} not built (python -m atmosphericturbulencesimul_amd.build)")
    lib = ctypes.CDLL(LIB_PATH)
    missing = []
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        for name in sorted(declared(h)):
            if not hasattr(lib, name):
                missing.append((os.path.basename(h), name))
    assert not missing, missing


def test_python_binding_covers_the_abi():
    from atmosphericturbulencesimul_amd import _lib
    decl = declared(os.path.join(ROOT, "include", "atsim.h"))
    assert decl <= set(_lib.exported_symbols()), decl - set(_lib.exported_symbols())
