"""CPU checks of the C-ABI boundary: the header parses as C, and the built
shared object loads and exports every function the header declares."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "parsec_hip.h")
LIB = os.path.join(ROOT, "parsec_python_amd", "libparsec_hip.so")


def declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\w[\w\s\*]*?\b(pc_\w+)\s*\(", text, flags=re.M)))


def test_header_is_valid_c():
    r = subprocess.run(["gcc", "-fsyntax-only", "-std=c99", "-x", "c", HEADER], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_library_exports_every_declared_symbol():
    if not os.path.exists(LIB):
        pytest.skip("libparsec_hip.so not built")
    lib = ctypes.CDLL(LIB)
    names = declared()
    assert len(names) > 40
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert lib.pc_version() == 1


def test_python_binding_covers_header():
    from parsec_python_amd import _lib
    proto = set(_lib.PROTOTYPES)
    assert set(declared()) <= proto
