"""The C ABI (include/goofit_hip.h): every declared entry point is exported by
the built library, with C linkage, and the ctypes wrapper names the same set."""
import ctypes
import os
import re
import subprocess

import goofit_amd
from goofit_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "goofit_hip.h")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gf_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared()
    assert "gf_reduce_metric" in names and "gf_dalitz_gram" in names
    assert len(names) >= 25


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(capi.LIB_PATH)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing
    # plain C names (no C++ mangling) in the dynamic symbol table
    out = subprocess.run(["nm", "-D", "--defined-only", capi.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    syms = {line.split()[-1] for line in out.splitlines() if line.strip()}
    assert set(declared()) <= syms


def test_ctypes_wrapper_matches_header():
    assert sorted(capi.EXPORTS) == declared()


def test_header_compiles_as_c():
    """The header is plain C (no C++ or HIP types cross the boundary)."""
    src = '#include "goofit_hip.h"\nint main(void) { gf_ctx *c = 0; (void)c; return gf_device_count() < 0; }\n'
    r = subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-fsyntax-only", "-I", os.path.dirname(HEADER),
                        "-x", "c", "-"], input=src, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_product_does_not_use_the_oracle():
    """Nothing under goofit_amd/ imports or links the oracle."""
    pkg = os.path.dirname(goofit_amd.__file__)
    for dirpath, _, files in os.walk(pkg):
        if "_build" in dirpath:
            continue
        for f in files:
            if f.endswith((".py", ".cpp", ".hip", ".h")):
                text = open(os.path.join(dirpath, f), errors="replace").read()
                assert not re.search(r"^\s*(import|from)\s+oracle\b", text, flags=re.M), f
                assert "liboracle" not in text and "goofit_oracle" not in text, f
