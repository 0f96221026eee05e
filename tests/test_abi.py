"""The C-ABI libraries load and export every function their headers declare
(no GPU call is made)."""
import ctypes
import os
import re

import fsig_amd as F
from conftest import ROOT

DECL = re.compile(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\**\s*\**\s*(fsig_[a-z0-9_]+)\s*\(", re.M)


def declared(header):
    with open(os.path.join(ROOT, "include", header)) as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(DECL.findall(text)))


def test_hip_library_exports_header():
    names = declared("fsig.h")
    assert "fsig_match" in names and "fsig_ctx_create" in names
    lib = ctypes.CDLL(F.HIP_LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_host_library_exports_header():
    names = declared("fsig_host.h")
    assert "fsig_hg_read" in names
    lib = ctypes.CDLL(F.HOST_LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_header():
    lib = F.hip_lib()  # binds argtypes; loading touches no GPU
    for n in declared("fsig.h"):
        assert getattr(lib, n).argtypes is not None or n in ("fsig_ctx_destroy",), n
