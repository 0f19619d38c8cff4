"""GPU compound batch (Compound::ComputeHydrodynamicForces, Compound.cpp:291-437)
against the FP64 oracle: an AUV-like hull with four thruster fairings,
flat sea and waves, states spread over outside / inside / crossing.

Bound per F/T vector: |x - x64| <= 1e-6 |x64| + 1e-7 S with S the compound's
natural scale (rho g V for buoyancy forces, rho u^3 A |C| for drag with
u = |v| + |omega| L; torques times the length L), since a compound sums parts
whose per-face scales the oracle does not report."""
import numpy as np
import pytest

import oracle as O
from stonefish_amd import transforms as T
from stonefish_amd import workloads as WL

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def log(msg):
    print(f"[compound] {msg}", flush=True)


def build(n, waves):
    from stonefish_amd import hydro as H
    from stonefish_amd import waves as W
    auv = WL.auv_compound()
    gmeshes = {}
    parts_g, parts_o = [], []
    omeshes = {}
    for p in auv.parts:
        key = id(p.mesh)
        if key not in gmeshes:
            gmeshes[key] = H.Mesh(p.mesh)
            omeshes[key] = O.Mesh.from_tri(p.mesh)
        parts_g.append(H.PartDesc(mesh=gmeshes[key], origin=p.origin, Cd=p.Cd, Cf=p.Cf, external=p.external,
                                  buoyant=p.buoyant))
        solid = O.make_body(omeshes[key], Cd=p.Cd, Cf=p.Cf, volume=p.mesh.volume(), surface=p.mesh.area(),
                            buoyant=p.buoyant)
        parts_o.append((solid, p.origin, T.identity_gl(), p.external))
    body = H.BodyDesc(mesh=None, aabb_min=auv.aabb_min, aabb_max=auv.aabb_max, volume=auv.volume,
                      surface=auv.surface, P_CB=auv.P_CB)
    batch = H.CompoundBatch([H.CompoundDesc(body, parts_g)] * n)
    dummy = O.Mesh(np.zeros((3, 3), np.float32), np.zeros((1, 3), np.uint32))
    ocomp = O.make_body(dummy, aabb_min=list(auv.aabb_min), aabb_max=list(auv.aabb_max), volume=auv.volume,
                        surface=auv.surface, P_CB=list(auv.P_CB))
    go = H.Ocean()
    oo = O.Ocean()
    grid = None
    if waves:
        comps = W.pm_components(32, 12.0, 256.0, seed=11)
        grid = torch.empty(256 * 256, dtype=torch.float32, device="cuda")
        W.build_grid_device(comps, 256, grid.data_ptr())
        torch.cuda.synchronize()
        g = grid.cpu().numpy()
        go.set_waves(g, 256, 1, [256.0])
        oo = O.Ocean(wave_data=g, wave_n=256, channels=1, lengths=[256.0])
    return auv, batch, go, oo, ocomp, parts_o, (gmeshes, omeshes, grid)


def states_for(rng, n, depth=(-0.8, 0.8)):
    from stonefish_amd import surrogate as S
    return S.sample_states(rng, n, xy_range=(0.0, 256.0), depth_range=depth)


@pytest.mark.parametrize("waves", [False, True])
def test_compound_batch_matches_oracle(gpu, waves):
    log(f"case: AUV compound batch, waves={waves}")
    n = 24
    auv, batch, go, oo, ocomp, parts_o, keep = build(n, waves)
    rng = np.random.default_rng(5 + int(waves))
    st = np.concatenate([states_for(rng, n - 6), states_for(rng, 3, (3.0, 4.0)), states_for(rng, 3, (-4.0, -3.0))])
    g = batch.compute_host(go, st)
    rho, grav = 1000.0, 9.81
    L = 3.0
    counts = np.zeros(3, int)
    for i in range(n):
        r = O.compound_forces(ocomp, parts_o, oo, st[i, :16], st[i, 16:19], st[i, 19:22], init=np.zeros(24))
        assert g[i, 21] == r[21] and g[i, 20] == r[20], i
        counts[int(r[21])] += 1
        u = np.linalg.norm(st[i, 16:19]) + np.linalg.norm(st[i, 19:22]) * L
        SF = {"b": rho * grav * auv.volume, "d": rho * u ** 3 * auv.surface}
        for j, name in enumerate(("Fb", "Tb", "Fdq", "Tdq", "Fdf", "Tdf")):
            if not (int(r[20]) >> j) & 1:
                continue
            sl = slice(3 * j, 3 * j + 3)
            S = SF["b" if j < 2 else "d"] * (L if j % 2 else 1.0)
            err = np.linalg.norm(g[i, sl] - r[sl])
            bound = 1e-6 * np.linalg.norm(r[sl]) + 1e-7 * S
            assert err <= bound, (i, name, err, bound, g[i, sl], r[sl])
        np.testing.assert_allclose(g[i, 18:20], r[18:20], rtol=1e-6, atol=1e-9)
    log(f"positions (outside, inside, crossing): {counts}")
    assert counts[1] > 0 and counts[2] > 0 and counts[0] > 0


def test_compound_masked_and_features(gpu):
    from stonefish_amd import surrogate as S
    n = 64
    auv, batch, go, oo, ocomp, parts_o, keep = build(n, True)
    rng = np.random.default_rng(21)
    st = states_for(rng, n)
    ds = torch.from_numpy(st).cuda()
    full = torch.zeros(n, 24, dtype=torch.float64, device="cuda")
    batch.compute(go, ds.data_ptr(), full.data_ptr())
    mask = torch.from_numpy((rng.random(n) < 0.4).astype(np.uint8)).cuda()
    part = torch.full((n, 24), -7.0, dtype=torch.float64, device="cuda")
    batch.compute_masked(go, ds.data_ptr(), mask.data_ptr(), part.data_ptr())
    torch.cuda.synchronize()
    m = mask.cpu().numpy().astype(bool)
    f, p = full.cpu().numpy(), part.cpu().numpy()
    assert np.array_equal(p[m, :22], f[m, :22]) and np.all(p[~m] == -7.0)
    # features use the compound's centre of buoyancy
    raw = torch.zeros(n, 13, dtype=torch.float32, device="cuda")
    S.features(batch, go, ds, 0, n, raw=raw)
    r = raw.cpu().numpy()
    import ctypes
    half = 0.5 * (auv.aabb_max - auv.aabb_min)
    wts = np.outer([0.25, 0.5, 0.25], [0.25, 0.5, 0.25])
    for i in range(0, n, 9):
        Rm = st[i, :16].reshape(4, 4).T[:3, :3]
        cb = Rm @ auv.P_CB + st[i, 12:15]
        # depth below the plane fitted over the footprint of the compound's AABB
        ex, ey = max(np.abs(Rm[0]) @ half, 1.0), max(np.abs(Rm[1]) @ half, 1.0)
        Hs = np.array([[-O.load().orc_wave_height_64(ctypes.byref(oo.s), cb[0] + (a - 1) * ex, cb[1] + (c - 1) * ey)
                        for c in range(3)] for a in range(3)])
        d = cb[2] + (wts * Hs).sum()
        assert abs(r[i, 9] - d) <= 2e-6 * (1 + abs(d))
