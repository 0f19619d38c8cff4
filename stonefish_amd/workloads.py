"""Synthetic workloads (seeded): bodies, states and sea states.

  capsule_submerged   capsule r 0.15 m, 1.6 m tip to tip, 2,048 triangles,
                      CG 5 m deep, flat sea: always fully submerged.
  ellipsoid_surface   ellipsoid half-axes (1.0, 0.3, 0.25), icosphere level 5
                      (20,480 triangles), CoB depth U[-0.2, 0.2]: always crossing.
  fleet               N bodies, each an ellipsoid or a box with half-axes
                      U[0.1, 2]^3 and a face count log-uniform in [500, 50,000],
                      CoB depth U[-1, 3], on a 1024 x 1024 m wave tile built
                      from 64 Pierson-Moskowitz sinusoids (U = 12 m/s).

Per step: linear velocity U[-2, 2]^3 m/s, angular velocity U[-1, 1]^3 rad/s,
orientation a uniform random unit quaternion.  Drag coefficients follow the
reference's ellipsoid approximation for a body of known half-axes
(SolidEntity.cpp:1014-1019): Cd_i = (1/a_i) / max_j(1/a_j), Cf = 0.1 Cd.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import meshes as M
from . import transforms as T
from . import waves as W


@dataclass
class BodyData:
    mesh: M.TriMesh
    half_axes: np.ndarray
    volume: float
    surface: float
    Cd: np.ndarray
    Cf: np.ndarray
    T_CG2C: np.ndarray = field(default_factory=T.identity_gl)
    T_CG2O: np.ndarray = field(default_factory=T.identity_gl)
    buoyant: bool = True


@dataclass
class SeaState:
    density: float = 1000.0
    gravity: np.ndarray = field(default_factory=lambda: np.array([0.0, 0.0, 9.81]))
    wave_components: np.ndarray | None = None   # pm_components() output
    wave_n: int = 0
    wave_len: float = 0.0


@dataclass
class Workload:
    name: str
    bodies: list[BodyData]
    sea: SeaState
    seed: int
    depth_range: tuple[float, float]
    xy_range: tuple[float, float]


def coeffs_from_half_axes(h) -> tuple[np.ndarray, np.ndarray]:
    cd = 1.0 / np.asarray(h, dtype=np.float64)
    cd /= cd.max()
    return cd, 0.1 * cd


def body_from_mesh(mesh: M.TriMesh, half_axes) -> BodyData:
    cd, cf = coeffs_from_half_axes(half_axes)
    return BodyData(mesh=mesh, half_axes=np.asarray(half_axes, dtype=np.float64), volume=mesh.volume(),
                    surface=mesh.area(), Cd=cd, Cf=cf)


def capsule_submerged(seed: int = 0) -> Workload:
    m = M.capsule(0.15, 1.6)
    b = body_from_mesh(m, (0.8, 0.15, 0.15))
    return Workload("capsule_submerged_2048", [b], SeaState(), seed, (5.0, 5.0), (0.0, 0.0))


def ellipsoid_surface(seed: int = 1) -> Workload:
    m = M.ico_ellipsoid(1.0, 0.3, 0.25, 5)
    b = body_from_mesh(m, (1.0, 0.3, 0.25))
    return Workload("ellipsoid_surface_20480", [b], SeaState(), seed, (-0.2, 0.2), (-50.0, 50.0))


def fleet(n_bodies: int = 4096, seed: int = 7, f_lo: int = 500, f_hi: int = 50000, waves: bool = True,
          wave_n: int = 1024, wave_len: float = 1024.0) -> Workload:
    rng = np.random.default_rng(seed)
    bodies = []
    kinds = rng.integers(0, 2, n_bodies)
    halfs = rng.uniform(0.1, 2.0, (n_bodies, 3))
    nfs = np.exp(rng.uniform(np.log(f_lo), np.log(f_hi), n_bodies))
    for i in range(n_bodies):
        h = halfs[i]
        if kinds[i] == 0:
            m = M.ellipsoid_with_faces(h[0], h[1], h[2], int(nfs[i]))
        else:
            m = M.box_with_faces(h[0], h[1], h[2], int(nfs[i]))
        bodies.append(body_from_mesh(m, h))
    sea = SeaState()
    if waves:
        sea.wave_components = W.pm_components(64, 12.0, wave_len, seed=seed + 1)
        sea.wave_n = wave_n
        sea.wave_len = wave_len
    return Workload(f"fleet_{n_bodies}", bodies, sea, seed, (-1.0, 3.0), (0.0, wave_len))


def random_states(w: Workload, steps: int, seed: int | None = None, positions: np.ndarray | None = None) -> np.ndarray:
    """(steps, nbodies, 22) float64: T_CG[16], v[3], omega[3] per body.  Body
    positions (x, y, depth) are drawn once per workload seed unless given."""
    nb = len(w.bodies)
    prng = np.random.default_rng(w.seed + 1000)
    if positions is None:
        positions = np.stack([prng.uniform(*w.xy_range, nb), prng.uniform(*w.xy_range, nb),
                              prng.uniform(*w.depth_range, nb)], axis=1)
    rng = np.random.default_rng(w.seed if seed is None else seed)
    out = np.zeros((steps, nb, 22))
    for s in range(steps):
        q = T.random_unit_quaternions(rng, nb)
        out[s, :, :16] = T.quat_to_gl(q, positions)
        out[s, :, 16:19] = rng.uniform(-2.0, 2.0, (nb, 3))
        out[s, :, 19:22] = rng.uniform(-1.0, 1.0, (nb, 3))
    return out


# ---- compound bodies: an AUV-like hull with four thruster fairings -------------------

@dataclass
class PartData:
    mesh: M.TriMesh
    origin: np.ndarray          # OpenGL 4x4: compound O frame -> part O frame
    Cd: np.ndarray
    Cf: np.ndarray
    external: bool = True
    buoyant: bool = True


@dataclass
class CompoundData:
    parts: list[PartData]
    volume: float
    surface: float
    P_CB: np.ndarray            # CG frame (the CG is the origin of the compound)
    aabb_min: np.ndarray        # CG-frame AABB of the collision shape
    aabb_max: np.ndarray
    mass: float
    inertia: np.ndarray         # principal moments about the CG axes


def auv_compound(hull_radius: float = 0.3, hull_length: float = 3.0, fairing_radius: float = 0.07,
                 fairing_length: float = 0.4, fairing_offset: float = 0.42, density: float = 1000.0,
                 buoyancy_margin: float = 0.02) -> CompoundData:
    """AUV-like compound: a capsule hull of 8,192 triangles along x and four
    capsule thruster fairings of 1,024 triangles each at the stern (at +-y and
    +-z).  The CG is at the origin; mass is the displaced water's minus a
    small margin (slightly positively buoyant)."""
    hull = M.capsule(hull_radius, hull_length, n_around=64, cap_bands=12, body_bands=41, axis=0)
    fair = M.capsule(fairing_radius, fairing_length, n_around=16, cap_bands=6, body_bands=21, axis=0)
    x_f = -0.5 * hull_length + hull_radius + 0.5 * fairing_length
    parts = [PartData(hull, T.identity_gl(), *coeffs_from_half_axes((0.5 * hull_length, hull_radius, hull_radius)))]
    for (y, z) in ((fairing_offset, 0.0), (-fairing_offset, 0.0), (0.0, fairing_offset), (0.0, -fairing_offset)):
        parts.append(PartData(fair, T.translation_gl((x_f, y, z)),
                              *coeffs_from_half_axes((0.5 * fairing_length, fairing_radius, fairing_radius))))
    vols = np.array([p.mesh.volume() for p in parts])
    cents = np.array([p.mesh.verts.astype(np.float64).mean(0) + p.origin[12:15] for p in parts])
    volume = float(vols.sum())
    P_CB = (vols[:, None] * cents).sum(0) / volume
    lo = np.min([p.mesh.verts.min(0) + p.origin[12:15] for p in parts], axis=0).astype(np.float64)
    hi = np.max([p.mesh.verts.max(0) + p.origin[12:15] for p in parts], axis=0).astype(np.float64)
    mass = density * volume * (1.0 - buoyancy_margin)
    r, L = hull_radius, hull_length
    inertia = np.array([0.5 * mass * r * r, mass * (3 * r * r + L * L) / 12.0, mass * (3 * r * r + L * L) / 12.0])
    return CompoundData(parts, volume, float(sum(p.mesh.area() for p in parts)), P_CB, lo, hi, mass, inertia)
