"""Edge cases of the batched path against the FP64 oracle: ragged and tiny
meshes (one triangle, a tetrahedron, a partly filled last chunk), a
zero-area face, vertices exactly on the surface, bodies that are not
hydrodynamic or not buoyant, every faces-per-lane setting, and an empty
batch.  Same bound as the other parity tests (tests/util.py)."""
import sys

import numpy as np
import pytest

import oracle as O
from stonefish_amd import hydro as H
from stonefish_amd import meshes as M
from stonefish_amd import transforms as T
from stonefish_amd import waves as W
from util import bound_ratio

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _meshes():
    tet = M.TriMesh(np.array([[0, 0, 0], [0.6, 0, 0], [0, 0.5, 0], [0, 0, 0.4]], np.float32) - 0.15,
                    np.array([[0, 2, 1], [0, 1, 3], [0, 3, 2], [1, 2, 3]], np.uint32))
    tri = M.TriMesh(np.array([[0, 0, 0], [0.5, 0, 0], [0, 0.5, 0]], np.float32),
                    np.array([[0, 1, 2]], np.uint32))
    box = M.box(0.3, 0.2, 0.1, 2)
    degenerate = M.TriMesh(box.verts, np.concatenate([box.faces, np.array([[0, 0, 1]], np.uint32)]))
    ragged = M.ico_ellipsoid(0.4, 0.3, 0.2, 1)         # 80 faces: a partly filled chunk at every K
    big = M.ico_ellipsoid(1.2, 0.5, 0.4, 4)            # 5,120 faces: several chunks
    return [("tetrahedron", tet), ("one triangle", tri), ("zero-area face", degenerate), ("80 faces", ragged),
            ("5120 faces", big), ("box", box)]


@pytest.mark.parametrize("waves", [False, True])
def test_edge_bodies_every_faces_per_lane(gpu, waves):
    log(f"case: edge-case bodies, waves={waves}")
    rng = np.random.default_rng(21 + int(waves))
    go, oo = H.Ocean(), None
    if waves:
        comps = W.pm_components(16, 6.0, 64.0, seed=5)
        grid = W.grid_reference(comps, 64).astype(np.float32)
        go.set_waves(grid, 64, 1, [64.0])
        oo = O.Ocean(wave_data=grid, wave_n=64, channels=1, lengths=[64.0])
    else:
        oo = O.Ocean()
    descs, obodies, states = [], [], []
    for k, (name, tm) in enumerate(_meshes()):
        for variant in range(3):
            buoyant = not (variant == 2 and k % 2)
            hydro = not (variant == 2 and k % 2 == 0)
            gm, om = H.Mesh(tm), O.Mesh.from_tri(tm)
            descs.append(H.BodyDesc(mesh=gm, volume=0.05, surface=0.5, buoyant=buoyant, hydro=hydro))
            obodies.append(O.make_body(om, volume=0.05, surface=0.5, buoyant=buoyant, hydro=hydro))
            s = np.zeros(22)
            if variant == 0 and not waves:
                # identity orientation with the mesh's top at depth exactly 0
                s[:16] = T.quat_to_gl(np.array([0.0, 0.0, 0.0, 1.0]), [1.0, 2.0, -float(tm.verts[:, 2].min())])
            else:
                depth = [0.0, rng.uniform(-0.3, 0.3), rng.uniform(-1, 2)][variant]
                s[:16] = T.quat_to_gl(T.random_unit_quaternions(rng, 1)[0], [rng.uniform(0, 60), rng.uniform(0, 60),
                                                                          depth])
            s[16:19] = rng.uniform(-2, 2, 3)
            s[19:22] = rng.uniform(-1, 1, 3)
            states.append(s)
    states = np.array(states)
    n = len(descs)
    init = np.full((n, 24), 0.375)
    rr = [O.body_forces(64, obodies[i], oo, states[i, :16], states[i, 16:19], states[i, 19:22], init=init[i],
                        full=True) for i in range(n)]
    r = np.stack([a for a, _ in rr])
    x = np.stack([b for _, b in rr])
    Cd = np.ones((n, 3))
    Cf = np.full((n, 3), 0.1)
    log(f"positions: {np.bincount((r[:, 21] + 1).astype(int), minlength=4)} (not hydro, outside, inside, crossing)")
    batch = H.BodyBatch(descs)
    for K in (1, 2, 4, 8, 16):
        batch.set_opts(faces_per_lane=K)
        g = batch.compute_host(go, states, init.copy())
        np.testing.assert_array_equal(g[:, 20:22], r[:, 20:22])
        untouched = r[:, 20] == 0
        np.testing.assert_array_equal(g[untouched, :20], init[untouched, :20])
        worst, _ = bound_ratio(g, r, x, Cd=Cd, Cf=Cf)
        log(f"  K={K}: " + ", ".join(f"{k} {v:.2f}" for k, v in worst.items()))
        assert max(worst.values()) <= 1.0


def test_empty_batch(gpu):
    batch = H.BodyBatch([])
    out = batch.compute_host(H.Ocean(), np.zeros((0, 22)))
    assert out.shape == (0, 24)
