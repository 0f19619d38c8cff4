"""The C++ host interface (include/sfhydro.hpp, reference call signatures)
drives the GPU path; its outputs against the FP64 oracle."""
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

import hull as HO
import oracle as O
from stonefish_amd import meshes as M
from stonefish_amd import transforms as T
from util import bound_ratio

pytestmark = pytest.mark.gpu
ROOT = Path(__file__).resolve().parents[1]
EXE = ROOT / "stonefish_amd" / "_lib" / "sfh_cpp_check"


def test_cpp_interface(gpu, tmp_path):
    print("case: C++ interface, 6 states", flush=True)
    assert EXE.exists(), "build the C++ check program first (__graft_entry__.build())"
    tm = M.ico_ellipsoid(0.9, 0.35, 0.3, 4)
    rng = np.random.default_rng(12)
    states = []
    for s in range(6):
        Tcg = T.quat_to_gl(T.random_unit_quaternions(rng, 1)[0], [rng.uniform(-20, 20), rng.uniform(-20, 20),
                                                                   rng.uniform(-0.25, 0.25) if s % 2 else 3.0])
        Tc = T.gl_mul(Tcg, T.translation_gl([0.02, -0.03, 0.01]))
        states.append(np.concatenate([Tcg, Tc, rng.uniform(-2, 2, 3), rng.uniform(-1, 1, 3)]))
    states = np.array(states)
    inp, outp = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(inp, "wb") as fh:
        np.array([tm.nv, tm.nf], dtype=np.uint32).tofile(fh)
        tm.verts.astype(np.float32).tofile(fh)
        tm.faces.astype(np.uint32).tofile(fh)
        np.array([len(states)], dtype=np.uint32).tofile(fh)
        states.astype(np.float64).tofile(fh)
    r = subprocess.run([str(EXE), str(inp), str(outp)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    out = np.fromfile(outp, dtype=np.float64).reshape(len(states), 56)
    om, oc = O.Mesh.from_tri(tm), O.Ocean()
    g, ref, ext = [], [], []
    gs, rs, xs = [], [], []
    for s, st in enumerate(states):
        a, e = O.surface(64, 1, 1, om, oc, st[:16], st[16:32], st[32:35], st[35:38], full=True)
        gg = np.zeros(24)
        gg[:20] = out[s, :20]
        g.append(gg); ref.append(a); ext.append(e)
        b, f = O.submerged(64, om, oc, st[:16], st[16:32], st[32:35], st[35:38], full=True)
        b[20] = 60
        h = np.zeros(24)
        h[6:18] = out[s, 20:32]
        gs.append(h); rs.append(b); xs.append(f)
        cor = O.correct(oc, out[s, 20:23], out[s, 23:26], out[s, 26:29], out[s, 29:32], [1.0, 0.5, 0.25],
                        [0.1, 0.05, 0.025], st[:16])
        np.testing.assert_allclose(out[s, 32:44], np.concatenate(cor), rtol=1e-13, atol=0)
        Fa, Ta, (SF, ST) = O.aero(om, 1.2, [(0, [2.0, -1.0, 0.5])], st[:16], st[16:32], st[32:35], st[35:38],
                                  sums=True)
        assert np.linalg.norm(out[s, 44:47] - Fa) <= 1e-6 * np.linalg.norm(Fa) + 1e-7 * SF
        assert np.linalg.norm(out[s, 47:50] - Ta) <= 1e-6 * np.linalg.norm(Ta) + 1e-7 * ST
        Fc, Tc_ = HO.correct_aero(HO.ELLIPSOID, [0.9, 0.35, 0.3], st[16:32], out[s, 44:47], out[s, 47:50])
        np.testing.assert_allclose(out[s, 50:56], np.concatenate([Fc, Tc_]), rtol=1e-13, atol=0)
    w1, p1 = bound_ratio(np.stack(g), np.stack(ref), np.stack(ext))
    w2, p2 = bound_ratio(np.stack(gs), np.stack(rs), np.stack(xs), names=("Fdq", "Tdq", "Fdf", "Tdf"))
    print("surface", w1, "submerged", w2, flush=True)
    assert max(w1.values()) <= 1.0 and max(w2.values()) <= 1.0
