"""GPU parity of the aerodynamic face loop (sfh_forces_aero, replacing
SolidEntity::ComputeAerodynamicForces, SolidEntity.cpp:1865-1918) against the
FP64 oracle, through the C ABI: every face of the mesh, normal-velocity drag
only, wind fields summed at each face centroid.

Bound per vector: |x - x64| <= 1e-6 |x64| + 1e-7 S, S the sum of the norms of
the per-face terms (the same bound as the hydrodynamic loops)."""
import sys

import numpy as np
import pytest

import oracle as O
from stonefish_amd import hydro as H
from stonefish_amd import meshes as M
from stonefish_amd import transforms as T

pytestmark = pytest.mark.gpu
RHO_AIR = 1.225


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _winds(centre):
    c = np.asarray(centre, dtype=float)
    jet = (c + [-2.0, 0.1, 0.05], [1.0, 0.05, 0.0], 0.4, 3.0)
    pipe = (c + [0.3, -3.0, -0.2], c + [0.3, 3.0, 0.2], 0.8, 0.5, 1.5, 1.5)
    atm = H.Atmosphere(RHO_AIR)
    atm.add_uniform_current([2.0, -1.0, 0.3])
    atm.add_jet(*jet)
    atm.add_pipe(*pipe)
    wind = [(0, [2.0, -1.0, 0.3])]
    d = np.asarray(jet[1]) / np.linalg.norm(jet[1])
    wind.append((1, list(jet[0]) + list(d) + [jet[2], jet[3]]))
    p1, p2, r1, r2, vin, gam = pipe
    ln = float(np.linalg.norm(p2 - p1))
    wind.append((2, list(p1) + list((p2 - p1) / ln) + [ln, r1, r2, vin, gam]))
    return atm, wind


@pytest.mark.parametrize("shape", ["capsule", "box", "ellipsoid"])
@pytest.mark.parametrize("with_wind", [False, True])
def test_aero_matches_oracle(gpu, shape, with_wind):
    log(f"case: aero {shape}, wind={with_wind}")
    tm = {"capsule": lambda: M.capsule(0.3, 2.0),
          "box": lambda: M.box(0.7, 0.4, 0.2, 8),
          "ellipsoid": lambda: M.ico_ellipsoid(1.0, 0.3, 0.25, 4)}[shape]()
    gm, om = H.Mesh(tm), O.Mesh.from_tri(tm)
    centre = [0.0, 0.0, -20.0]
    if with_wind:
        atm, wind = _winds(centre)
    else:
        atm, wind = H.Atmosphere(RHO_AIR), []
    rng = np.random.default_rng(31 + len(shape) + int(with_wind))
    worst_F = worst_T = 0.0
    for s in range(12):
        q = T.random_unit_quaternions(rng, 1)[0]
        Tcg = T.quat_to_gl(q, np.asarray(centre) + rng.uniform(-0.3, 0.3, 3))
        Tc = T.quat_to_gl(q, Tcg[12:15] + rng.uniform(-0.05, 0.05, 3)) if s % 2 else Tcg
        v, w = rng.uniform(-8, 8, 3), rng.uniform(-2, 2, 3)
        F, Tq = np.zeros(3), np.zeros(3)
        H.ComputeAerodynamicForces(gm, atm, Tcg, Tc, v, w, F, Tq)
        Fr, Tr, (SF, ST) = O.aero(om, RHO_AIR, wind, Tcg, Tc, v, w, sums=True)
        assert np.linalg.norm(Fr) > 0
        worst_F = max(worst_F, np.linalg.norm(F - Fr) / (1e-6 * np.linalg.norm(Fr) + 1e-7 * SF))
        worst_T = max(worst_T, np.linalg.norm(Tq - Tr) / (1e-6 * np.linalg.norm(Tr) + 1e-7 * ST))
    log(f"  bound ratio F {worst_F:.2f}, T {worst_T:.2f}")
    assert worst_F <= 1.0 and worst_T <= 1.0


def test_aero_ignores_the_water_line_and_null_mesh(gpu):
    """The aerodynamic loop takes every face wherever the body is (the caller
    decides which bodies are in the air); a null mesh gives zeros."""
    tm = M.ico_ellipsoid(1.0, 0.3, 0.25, 3)
    gm, om = H.Mesh(tm), O.Mesh.from_tri(tm)
    atm = H.Atmosphere(RHO_AIR)
    rng = np.random.default_rng(4)
    for z in (-5.0, 0.0, 5.0):
        q = T.random_unit_quaternions(rng, 1)[0]
        Tcg = T.quat_to_gl(q, [0.1, -0.2, z])
        v, w = rng.uniform(-5, 5, 3), rng.uniform(-1, 1, 3)
        F, Tq = np.zeros(3), np.zeros(3)
        H.ComputeAerodynamicForces(gm, atm, Tcg, Tcg, v, w, F, Tq)
        Fr, Tr, (SF, ST) = O.aero(om, RHO_AIR, [], Tcg, Tcg, v, w, sums=True)
        assert np.linalg.norm(F - Fr) <= 1e-6 * np.linalg.norm(Fr) + 1e-7 * SF
        assert np.linalg.norm(Tq - Tr) <= 1e-6 * np.linalg.norm(Tr) + 1e-7 * ST
    F, Tq = np.full(3, 7.0), np.full(3, 7.0)
    H.ComputeAerodynamicForces(None, atm, Tcg, Tcg, v, w, F, Tq)
    assert not F.any() and not Tq.any()


# ---- batches: SolidEntity::ComputeAerodynamicForces(atm) and the compound version

import hull as HO   # noqa: E402  (FP64 restatement of CorrectAerodynamicForces)


def _gl_mul(a, b):
    A = np.asarray(a, dtype=float).reshape(4, 4).T
    B = np.asarray(b, dtype=float).reshape(4, 4).T
    return (A @ B).T.reshape(-1)


def _pose(rng, scale=1.0):
    return T.quat_to_gl(T.random_unit_quaternions(rng, 1)[0], rng.uniform(-scale, scale, 3))


def _check_corrected(g, Fr, Tr, SF, ST, kind, params, T_CG2H, what):
    """Corrected outputs against the oracle: the factor |Cd . d| (|Cd| <= 1.5)
    moves by at most 1.5 |dF| / |F| when F moves by dF, so the raw bounds bF,
    bT become cor bF + 1.5 bF and cor bT + 1.5 |T| bF / |F|."""
    Fc, Tc = HO.correct_aero(kind, params, T_CG2H, Fr, Tr)
    Fc, Tc = np.array(Fc), np.array(Tc)
    cor = np.linalg.norm(Fc) / np.linalg.norm(Fr)
    bF = 1e-6 * np.linalg.norm(Fr) + 1e-7 * SF
    bT = 1e-6 * np.linalg.norm(Tr) + 1e-7 * ST
    rF = np.linalg.norm(g[6:9] - Fc) / ((cor + 1.5) * bF)
    rT = np.linalg.norm(g[9:12] - Tc) / (cor * bT + 1.5 * np.linalg.norm(Tr) * bF / np.linalg.norm(Fr))
    assert rF <= 1.0 and rT <= 1.0, (what, rF, rT)
    return max(rF, rT)


def test_aero_batch_matches_oracle(gpu):
    """Bodies of several meshes and proxies (sphere, cylinder, ellipsoid; a
    body without aerodynamic mode, a body without mesh) in a wind: each row
    equals the oracle's aerodynamic integral after CorrectAerodynamicForces;
    the non-aerodynamic row is untouched."""
    log("case: aero batch, 5 shapes x 3, wind, graph replay")
    rng = np.random.default_rng(41)
    atm, wind = _winds([0.0, 0.0, -20.0])
    shapes = [M.capsule(0.3, 2.0), M.box(0.7, 0.4, 0.2, 8), M.ico_ellipsoid(1.0, 0.3, 0.25, 4),
              M.box_with_faces(0.3, 0.5, 0.4, 3000), M.ico_ellipsoid(0.4, 0.4, 0.4, 3)]
    descs, refs = [], []
    for i in range(15):
        tm = shapes[i % len(shapes)]
        kind = [HO.SPHERE, HO.CYLINDER, HO.ELLIPSOID][i % 3]
        params = rng.uniform(0.2, 1.5, 3)
        d = H.AeroDesc(H.Mesh(tm), T_CG2C=_pose(rng, 0.05), T_CG2H=_pose(rng, 0.1), approx_params=params,
                       approx_type=kind, aero=(i != 7))
        if i == 11:
            d.mesh = None
        descs.append(d)
        refs.append(O.Mesh.from_tri(tm))
    batch = H.AeroBatch(descs)
    n = len(descs)
    states = np.zeros((n, 22))
    for i in range(n):
        states[i, :16] = T.quat_to_gl(T.random_unit_quaternions(rng, 1)[0], [0, 0, -20.0] + rng.uniform(-0.3, 0.3, 3))
        states[i, 16:19] = rng.uniform(-8, 8, 3)
        states[i, 19:22] = rng.uniform(-2, 2, 3)
    out0 = np.full((n, 24), 3.25)
    out = batch.compute_host(atm, states, out0.copy())
    worst = 0.0
    for i, d in enumerate(descs):
        if not d.aero:
            assert out[i, 20] == 0 and out[i, 21] == -1
            np.testing.assert_array_equal(out[i, :20], out0[i, :20])
            continue
        assert out[i, 20] == 12 and out[i, 21] == 1
        if d.mesh is None:
            assert not out[i, 6:12].any()
            continue
        Tcg = states[i, :16]
        Fr, Tr, (SF, ST) = O.aero(refs[i], RHO_AIR, wind, Tcg, _gl_mul(Tcg, d.T_CG2C), states[i, 16:19],
                                  states[i, 19:22], sums=True)
        worst = max(worst, _check_corrected(out[i], Fr, Tr, SF, ST, d.approx_type, d.approx_params, d.T_CG2H,
                                            f"body {i}"))
    log(f"  worst bound ratio {worst:.2f}")
    # the same through the device entry points and a captured graph
    import torch
    dev = torch.device("cuda", 0)
    st = torch.from_numpy(states).to(dev)
    do = torch.from_numpy(out0).to(dev)
    s = torch.cuda.Stream()
    batch.capture(atm, st.data_ptr(), do.data_ptr(), s.cuda_stream)
    for _ in range(3):
        batch.replay(s.cuda_stream)
        s.synchronize()
        np.testing.assert_array_equal(do.cpu().numpy(), out)


def test_aero_compound_batch_matches_oracle(gpu):
    """Compounds of three parts (one internal) with their own origins, O2C
    frames and proxies: the row is the sum over external parts of the
    oracle's corrected part drag; a compound without aerodynamic mode keeps
    its row."""
    log("case: aero compound batch")
    rng = np.random.default_rng(43)
    atm, wind = _winds([0.0, 0.0, -20.0])
    meshes = [M.capsule(0.2, 1.5), M.box(0.3, 0.2, 0.1, 6), M.ico_ellipsoid(0.3, 0.2, 0.2, 3)]
    comps = []
    for c in range(6):
        parts = []
        for p in range(3):
            parts.append(H.AeroPartDesc(H.Mesh(meshes[p]), origin=_pose(rng, 0.5), O2C=_pose(rng, 0.05),
                                        T_CG2H=_pose(rng, 0.1), approx_params=rng.uniform(0.2, 1.0, 3),
                                        approx_type=int(rng.integers(0, 4)), external=(p != 1 or c % 2 == 0)))
        comps.append(H.AeroCompoundDesc(T_CG2O=_pose(rng, 0.1), parts=parts, aero=(c != 4)))
    batch = H.AeroCompoundBatch(comps)
    n = len(comps)
    states = np.zeros((n, 22))
    for i in range(n):
        states[i, :16] = T.quat_to_gl(T.random_unit_quaternions(rng, 1)[0], [0, 0, -20.0] + rng.uniform(-0.3, 0.3, 3))
        states[i, 16:19] = rng.uniform(-8, 8, 3)
        states[i, 19:22] = rng.uniform(-2, 2, 3)
    out0 = np.full((n, 24), -1.5)
    out = batch.compute_host(atm, states, out0.copy())
    oms = [O.Mesh.from_tri(m) for m in meshes]
    for i, c in enumerate(comps):
        if not c.aero:
            assert out[i, 20] == 0 and out[i, 21] == -1
            np.testing.assert_array_equal(out[i, :20], out0[i, :20])
            continue
        assert out[i, 20] == 12 and out[i, 21] == 1
        Tcg = states[i, :16]
        F, Tq, bF, bT = np.zeros(3), np.zeros(3), 0.0, 0.0
        for p, part in enumerate(c.parts):
            if not part.external:
                continue
            Tc = _gl_mul(_gl_mul(_gl_mul(Tcg, c.T_CG2O), part.origin), part.O2C)
            Fr, Tr, (SF, ST) = O.aero(oms[p], RHO_AIR, wind, Tcg, Tc, states[i, 16:19], states[i, 19:22], sums=True)
            Fc, Tc_ = HO.correct_aero(part.approx_type, part.approx_params, part.T_CG2H, Fr, Tr)
            cor = np.linalg.norm(Fc) / np.linalg.norm(Fr)
            F += Fc
            Tq += Tc_
            b1 = 1e-6 * np.linalg.norm(Fr) + 1e-7 * SF
            bF += (cor + 1.5) * b1
            bT += cor * (1e-6 * np.linalg.norm(Tr) + 1e-7 * ST) + 1.5 * np.linalg.norm(Tr) * b1 / np.linalg.norm(Fr)
        assert np.linalg.norm(out[i, 6:9] - F) <= bF, i
        assert np.linalg.norm(out[i, 9:12] - Tq) <= bT, i


def test_aero_batches_are_not_surrogate_batches(gpu):
    from stonefish_amd import _native as N
    batch = H.AeroBatch([H.AeroDesc(H.Mesh(M.box(0.3, 0.2, 0.1, 2)))])
    with pytest.raises(N.SfhError):
        N.check(N.load().sfh_surrogate_set_table(batch.handle, 0, 1, None, 0, 0, 0))
