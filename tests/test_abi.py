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


def test_struct_layouts_match_header(tmp_path):
    """Every ctypes mirror of a struct in include/sfhydro.h has the C
    compiler's size and field offsets (a field added on one side only, or a
    padding difference, would shift every later field)."""
    import shutil
    import subprocess
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    structs = [getattr(N, n) for n in sorted(dir(N)) if n.startswith("sfh_")
               and isinstance(getattr(N, n), type) and issubclass(getattr(N, n), C.Structure)]
    assert len(structs) >= 10
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "sfhydro.h"', "int main(void) {"]
    for cls in structs:
        nm = cls.__name__
        lines.append(f'printf("{nm} %zu\\n", sizeof({nm}));')
        for f in cls._fields_:
            lines.append(f'printf("{nm}.{f[0]} %zu\\n", offsetof({nm}, {f[0]}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run([cc, "-std=c11", "-I", str(ROOT / "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.splitlines()
    got = dict(line.rsplit(" ", 1) for line in out if line)
    for cls in structs:
        assert int(got[cls.__name__]) == C.sizeof(cls), cls.__name__
        for f in cls._fields_:
            assert int(got[f"{cls.__name__}.{f[0]}"]) == getattr(cls, f[0]).offset, (cls.__name__, f[0])
