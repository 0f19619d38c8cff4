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
