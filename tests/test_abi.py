"""The C-ABI library loads without a GPU and exports every entry point that
include/sfhydro.h declares (and the ctypes binding covers them all)."""
import ctypes as C
import re
from pathlib import Path

from stonefish_amd import _native as N

ROOT = Path(__file__).resolve().parents[1]


def declared(header: str):
    text = (ROOT / "include" / header).read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(sfh_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(str(N.lib_path()))
    names = declared("sfhydro.h")
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_header():
    assert set(declared("sfhydro.h")) == set(N.EXPORTED)


def test_last_error_without_device():
    lib = N.load()
    assert lib.sfh_last_error() is not None
    # argument validation needs no GPU and reports through the error string
    rc = lib.sfh_mesh_create(None, 12, 3, None, 1, C.byref(C.c_void_p()))
    assert rc == N.SFH_ERR_INVALID
    assert b"null" in lib.sfh_last_error()
