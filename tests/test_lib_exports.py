"""The C-ABI library loads without a GPU and exports every symbol that include/*.h declares (no compute calls)."""
import ctypes
import glob
import os
import re

from tropidist_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for path in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(path).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(tdp_\w+)\s*\(", text))
    return names


def test_library_exports_every_declared_symbol():
    assert os.path.exists(_lib.LIB_PATH), "build libtdp.so first (python -m tropidist_amd.build)"
    lib = ctypes.CDLL(_lib.LIB_PATH)
    names = declared_symbols()
    assert len(names) >= 10
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


def test_bindings_cover_the_header():
    """Every declared entry point has a ctypes signature in the Python binding."""
    assert declared_symbols() <= set(_lib.SIGNATURES)


def test_version_string_without_a_device():
    assert b"gfx950" in _lib.lib().tdp_version()
