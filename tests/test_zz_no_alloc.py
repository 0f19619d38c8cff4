"""sfh_forces_noalloc (the per-body evaluation that allocates nothing) against
the entry points it is a variant of, sfh_forces_surface and
sfh_forces_submerged, through the C ABI: the same kernels on the same chunks,
so the outputs must be bit-identical.  Two calls in a row, the second with a
smaller mesh (its work buffers are the first call's, reused), then a mesh
above the size limit, which must be refused before anything runs."""
import sys

import numpy as np
import pytest

from stonefish_amd import _native as N
from stonefish_amd import hydro as H
from stonefish_amd import meshes as M
from stonefish_amd import waves as W
from stonefish_amd import workloads as WL

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _static(settings, mode, gm, go, st):
    """The existing entry point: host outputs, untouched slots from out0."""
    outs = [np.zeros(3) for _ in range(6)]
    sv = np.zeros(2)
    if mode == N.SFH_NOALLOC_SURFACE:
        H.ComputeHydrodynamicForcesSurface(settings, gm, go, st[:16], st[16:32], st[32:35], st[35:38], *outs, sv)
    else:
        H.ComputeHydrodynamicForcesSubmerged(gm, go, st[:16], st[16:32], st[32:35], st[35:38], *outs[2:])
    return np.concatenate(outs + [sv])


def _noalloc(torch, settings, mode, gm, go, st):
    dev = torch.device("cuda:0")
    d_state = torch.tensor(st, dtype=torch.float64, device=dev)
    d_out = torch.zeros(24, dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream()
    H.forces_noalloc(settings, mode, gm, go, d_state.data_ptr(), d_out.data_ptr(), stream.cuda_stream)
    torch.cuda.synchronize()
    return d_out.cpu().numpy()


def test_noalloc_matches_the_per_body_entry_points(gpu):
    import torch
    go = H.Ocean()
    comps = W.pm_components(16, 6.0, 64.0, seed=3)
    grid = W.grid_reference(comps, 64).astype(np.float32)
    go.set_waves(grid, 64, 1, [64.0])
    settings = H.HydrodynamicsSettings(True, True)
    cases = [
        ("surface ellipsoid 20,480 faces, waves", WL.ellipsoid_surface(), N.SFH_NOALLOC_SURFACE),
        ("submerged capsule 2,048 faces (smaller, second call)", WL.capsule_submerged(), N.SFH_NOALLOC_SUBMERGED),
        ("surface box 48 faces (smaller again)", None, N.SFH_NOALLOC_SURFACE),
    ]
    rng = np.random.default_rng(5)
    for what, w, mode in cases:
        log(f"case: {what}")
        if w is None:
            tm = M.box(0.5, 0.3, 0.2, 2)
            T = np.eye(4)
            T[2, 3] = 0.1
            st = np.concatenate([T.T.reshape(-1), T.T.reshape(-1), rng.uniform(-2, 2, 3), rng.uniform(-1, 1, 3)])
        else:
            tm = w.bodies[0].mesh
            s = WL.random_states(w, 1)[0, 0]
            st = np.concatenate([s[:16], s[:16], s[16:19], s[19:22]])
        gm = H.Mesh(tm)
        ref = _static(settings, mode, gm, go, st)
        got = _noalloc(torch, settings, mode, gm, go, st)
        mask = int(got[20])
        log(f"  position {got[21]:.0f}, written mask {mask}")
        assert got[21] in (1.0, 2.0)
        for j in range(8):
            if (mask >> j) & 1:
                sl = slice(3 * j, 3 * j + 3) if j < 6 else slice(18 + j - 6, 19 + j - 6)
                np.testing.assert_array_equal(got[sl], ref[sl], err_msg=f"{what}: slot {j}")
        # slots the call does not write keep the caller's contents (zeros here)
        for j in range(6):
            if not (mask >> j) & 1:
                np.testing.assert_array_equal(got[3 * j:3 * j + 3], 0.0)
        if w is not None:
            assert np.abs(got[:18]).max() > 0.0


def test_noalloc_refuses_meshes_above_its_limit(gpu):
    """600,000 faces (above the 524,288-face limit): refused with SFH_ERR_SIZE
    before any launch, the output row untouched."""
    import ctypes as C

    import torch
    lib = N.load()
    verts = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    faces = np.tile(np.array([[0, 1, 2]], np.uint32), (600_000, 1))
    gm = H.Mesh(M.TriMesh(verts, faces))
    go = H.Ocean()
    dev = torch.device("cuda:0")
    d_state = torch.zeros(38, dtype=torch.float64, device=dev)
    d_out = torch.zeros(24, dtype=torch.float64, device=dev)
    s = N.sfh_settings(1, 1)
    rc = lib.sfh_forces_noalloc(C.byref(s), N.SFH_NOALLOC_SURFACE, gm.handle, go.handle,
                                C.c_void_p(d_state.data_ptr()), C.c_void_p(d_out.data_ptr()), None)
    assert rc == N.SFH_ERR_SIZE
    assert b"no-alloc limit" in lib.sfh_last_error()
    torch.cuda.synchronize()
    assert float(d_out.abs().max()) == 0.0
