"""Procedural closed triangle meshes (outward winding) and an OBJ reader.

Faces are wound counter-clockwise seen from outside, so (p2-p1)x(p3-p1) points
out of the body as the reference's per-face loops expect
(SolidEntity.cpp:1385-1389).  Faces are emitted band by band so neighbouring
faces share vertices that are close in memory.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass
class TriMesh:
    verts: np.ndarray   # (nv, 3) float32
    faces: np.ndarray   # (nf, 3) uint32

    @property
    def nv(self) -> int:
        return int(self.verts.shape[0])

    @property
    def nf(self) -> int:
        return int(self.faces.shape[0])

    def aabb(self):
        return self.verts.min(axis=0).astype(np.float64), self.verts.max(axis=0).astype(np.float64)

    def volume(self) -> float:
        """Enclosed volume (divergence theorem), float64."""
        v = self.verts.astype(np.float64)
        a, b, c = v[self.faces[:, 0]], v[self.faces[:, 1]], v[self.faces[:, 2]]
        return float(np.einsum("ij,ij->i", a, np.cross(b, c)).sum() / 6.0)

    def area(self) -> float:
        v = self.verts.astype(np.float64)
        a, b, c = v[self.faces[:, 0]], v[self.faces[:, 1]], v[self.faces[:, 2]]
        return float(0.5 * np.linalg.norm(np.cross(b - a, c - a), axis=1).sum())


def _revolve(profile_r: np.ndarray, profile_z: np.ndarray, n_around: int) -> TriMesh:
    """Closed surface of revolution about z.  profile runs from the bottom pole
    (r=0) to the top pole (r=0); interior rings have r > 0."""
    rings = len(profile_r) - 2
    phi = 2.0 * np.pi * np.arange(n_around) / n_around
    verts = [np.array([[0.0, 0.0, profile_z[0]]])]
    for k in range(1, rings + 1):
        r, z = profile_r[k], profile_z[k]
        verts.append(np.stack([r * np.cos(phi), r * np.sin(phi), np.full(n_around, z)], axis=1))
    verts.append(np.array([[0.0, 0.0, profile_z[-1]]]))
    V = np.concatenate(verts).astype(np.float32)
    top = V.shape[0] - 1
    faces = []
    j = np.arange(n_around)
    jn = (j + 1) % n_around
    ring = lambda k: 1 + (k - 1) * n_around  # noqa: E731  first vertex of ring k
    # bottom fan: pole, ring1[j+1], ring1[j]  (outward normal points to -z)
    faces.append(np.stack([np.zeros(n_around, int), ring(1) + jn, ring(1) + j], axis=1))
    for k in range(1, rings):
        a, b = ring(k) + j, ring(k) + jn
        c, d = ring(k + 1) + j, ring(k + 1) + jn
        faces.append(np.stack([a, b, d], axis=1))
        faces.append(np.stack([a, d, c], axis=1))
    faces.append(np.stack([ring(rings) + j, ring(rings) + jn, np.full(n_around, top)], axis=1))
    F = np.concatenate(faces).astype(np.uint32)
    return TriMesh(V, F)


def cylinder(radius: float, length: float, n_around: int = 32, bands: int = 4) -> TriMesh:
    """Closed cylinder about z, `length` end to end, flat caps (fans from the
    axis)."""
    zs = np.linspace(-0.5 * length, 0.5 * length, bands + 1)
    pr = np.concatenate([[0.0], np.full(bands + 1, radius), [0.0]])
    pz = np.concatenate([[zs[0]], zs, [zs[-1]]])
    return _revolve(pr, pz, n_around)


def capsule(radius: float, length: float, n_around: int = 32, cap_bands: int = 8, body_bands: int = 17,
            axis: int = 0) -> TriMesh:
    """Capsule, `length` tip to tip along `axis`.  Faces: 2*n_around*(2*cap_bands-1+body_bands).
    The default (32, 8, 17) gives 2,048 triangles and 1,026 vertices."""
    half_cyl = 0.5 * length - radius
    # profile from bottom pole to top pole
    th = np.linspace(-0.5 * np.pi, 0.0, cap_bands + 1)          # bottom cap, pole .. equator
    r_b, z_b = radius * np.cos(th), -half_cyl + radius * np.sin(th)
    zc = np.linspace(-half_cyl, half_cyl, body_bands + 1)[1:-1]  # cylinder interior rings
    r_c = np.full_like(zc, radius)
    th2 = np.linspace(0.0, 0.5 * np.pi, cap_bands + 1)
    r_t, z_t = radius * np.cos(th2), half_cyl + radius * np.sin(th2)
    pr = np.concatenate([r_b, r_c, r_t])
    pz = np.concatenate([z_b, zc, z_t])
    pr[0] = pr[-1] = 0.0
    m = _revolve(pr, pz, n_around)
    if axis != 2:
        perm = {0: [2, 0, 1], 1: [1, 2, 0]}[axis]   # move z onto the requested axis
        m = TriMesh(np.ascontiguousarray(m.verts[:, perm]), m.faces)
    return m


def uv_ellipsoid(a: float, b: float, c: float, n_around: int, n_rings: int) -> TriMesh:
    """Ellipsoid with half-axes (a, b, c) as a latitude/longitude mesh:
    2*n_around*(n_rings-1) faces."""
    th = np.linspace(-0.5 * np.pi, 0.5 * np.pi, n_rings + 1)
    m = _revolve(np.cos(th), np.sin(th), n_around)
    m.verts[:, 0] *= a
    m.verts[:, 1] *= b
    m.verts[:, 2] *= c
    m.verts[0] = (0, 0, -c)
    m.verts[-1] = (0, 0, c)
    return m


def ellipsoid_with_faces(a: float, b: float, c: float, target_faces: int) -> TriMesh:
    """UV ellipsoid whose face count is close to target_faces."""
    rings = max(3, int(round(np.sqrt(target_faces / 4.0))) + 1)
    around = max(3, int(round(target_faces / (2.0 * (rings - 1)))))
    return uv_ellipsoid(a, b, c, around, rings)


def icosphere(subdivisions: int) -> TriMesh:
    """Unit icosphere: 20*4**subdivisions faces, vertices on the unit sphere."""
    t = (1.0 + 5 ** 0.5) / 2.0
    V = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    F = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
         (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11),
         (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    verts = [np.array(v, dtype=np.float64) / np.linalg.norm(v) for v in V]
    faces = list(F)
    for _ in range(subdivisions):
        cache = {}

        def mid(i, j):
            key = (min(i, j), max(i, j))
            if key not in cache:
                p = verts[i] + verts[j]
                verts.append(p / np.linalg.norm(p))
                cache[key] = len(verts) - 1
            return cache[key]

        nf = []
        for (i, j, k) in faces:
            a, b, c = mid(i, j), mid(j, k), mid(k, i)
            nf += [(i, a, c), (j, b, a), (k, c, b), (a, b, c)]
        faces = nf
    return TriMesh(np.array(verts, dtype=np.float32), np.array(faces, dtype=np.uint32))


def ico_ellipsoid(a: float, b: float, c: float, subdivisions: int) -> TriMesh:
    m = icosphere(subdivisions)
    v = m.verts.astype(np.float64) * np.array([a, b, c])
    return TriMesh(v.astype(np.float32), m.faces)


def box(hx: float, hy: float, hz: float, n: int) -> TriMesh:
    """Axis-aligned box with each side split n x n: 12*n*n faces.  Each side has
    its own (n+1)^2 vertices, so edge vertices are duplicated across sides; the
    surface is geometrically closed."""
    g = np.linspace(-1.0, 1.0, n + 1)
    U, V = np.meshgrid(g, g, indexing="ij")            # (n+1, n+1), index [i, j]
    h = np.array([hx, hy, hz])
    verts, faces = [], []
    base = 0
    ii, jj = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    a = (ii * (n + 1) + jj).ravel()
    b = ((ii + 1) * (n + 1) + jj).ravel()
    c = ((ii + 1) * (n + 1) + jj + 1).ravel()
    d = (ii * (n + 1) + jj + 1).ravel()
    for axis in range(3):
        u_ax, v_ax = [(1, 2), (2, 0), (0, 1)][axis]
        for sgn in (-1.0, 1.0):
            p = np.zeros(((n + 1) * (n + 1), 3))
            p[:, axis] = sgn
            p[:, u_ax] = U.ravel()
            p[:, v_ax] = V.ravel()
            verts.append(p * h)
            if sgn > 0:
                f = np.concatenate([np.stack([a, b, c], 1), np.stack([a, c, d], 1)])
            else:
                f = np.concatenate([np.stack([a, c, b], 1), np.stack([a, d, c], 1)])
            # interleave the two triangles of each quad for locality
            f = f.reshape(2, -1, 3).transpose(1, 0, 2).reshape(-1, 3)
            faces.append(f + base)
            base += p.shape[0]
    return TriMesh(np.concatenate(verts).astype(np.float32), np.concatenate(faces).astype(np.uint32))


def box_with_faces(hx: float, hy: float, hz: float, target_faces: int) -> TriMesh:
    n = max(1, int(round(np.sqrt(target_faces / 12.0))))
    return box(hx, hy, hz, n)


def load_obj(path) -> TriMesh:
    """Minimal Wavefront OBJ reader: `v` lines and triangular or polygonal `f`
    lines (fan-triangulated), 1-based indices, texture/normal indices ignored."""
    V, F = [], []
    with open(path) as fh:
        for line in fh:
            if line.startswith("v "):
                V.append([float(x) for x in line.split()[1:4]])
            elif line.startswith("f "):
                idx = [int(tok.split("/")[0]) - 1 for tok in line.split()[1:]]
                for k in range(1, len(idx) - 1):
                    F.append((idx[0], idx[k], idx[k + 1]))
    return TriMesh(np.array(V, dtype=np.float32), np.array(F, dtype=np.uint32))
