"""The C-ABI libraries load without a GPU and export every entry point that
include/sfhydro.h and include/sfmlp.h declare (and the ctypes bindings cover
them all)."""
import ctypes as C
import re
from pathlib import Path

import pytest

from stonefish_amd import _native as N

ROOT = Path(__file__).resolve().parents[1]


def declared(header: str, prefix: str = "sfh"):
    text = (ROOT / "include" / header).read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(" + prefix + r"_[a-z0-9_]+)\s*\(", text)))


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


def test_mlp_library_exports_every_declared_symbol():
    lib = C.CDLL(str(N.mlp_lib_path()))
    names = declared("sfmlp.h", "sfm")
    assert len(names) >= 12
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(names) == set(N.MLP_EXPORTED)


def test_mlp_argument_checks_without_device():
    lib = N.load_mlp()
    rc = lib.sfm_create(100, 1, C.byref(C.c_void_p()))   # not a multiple of 256
    assert rc == -2
    assert b"multiple" in lib.sfm_last_error()
    assert lib.sfm_destroy(None) == -1


def test_current_argument_checks_without_device():
    """sfh_ocean_add_current: jets and pipes in the reference's parameter
    layout; bad types, non-unit axes, non-positive pipe lengths and a fifth
    jet or pipe are refused (host-side checks, no GPU needed)."""
    from stonefish_amd import hydro as H
    o = H.Ocean()
    o.add_uniform_current([0.2, 0.0, 0.0])
    o.add_jet([0, 0, 1], [2, 0, 0], 0.5, 3.0)   # the axis is normalised as in Jet::Jet
    o.add_pipe([0, 0, 0], [0, 4, 0], 1.0, 0.5, 1.0, 1 / 7)
    lib = N.load()
    for kind, params in ((1, [0, 0, 0, 1, 1, 0, 0.5, 1]), (3, [0] * 12), (2, [0, 0, 0, 1, 0, 0, 0.0, 1, 1, 1, 1])):
        with pytest.raises(N.SfhError):
            o.add_current(kind, params)
    o.add_jet([1, 0, 1], [0, 0, 1], 0.2, 1.0)
    o.add_pipe([0, 1, 0], [0, 1, 3], 0.4, 0.4, 1.0, 0.5)
    with pytest.raises(N.SfhError):
        o.add_jet([2, 0, 1], [0, 1, 0], 0.2, 1.0)
    assert b"at most 4" in lib.sfh_last_error()
    o.close()
