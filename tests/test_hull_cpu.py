"""Hull preprocessing (host): the library's sfh_compute_physical_properties and
sfh_compute_fluid_dynamics_approx against the FP64 restatement oracle/hull.py,
and the restatement itself pinned by the reference test's known answer for
hull_hydro.obj (volume and principal inertia, FluidDynamicsTestManager.cpp:
206-215), closed forms (boxes, rotated boxes, cylinders, a thin spherical
shell), the minimum-volume-ellipsoid iteration on ellipsoids, and Bullet's
Euler convention.  No GPU."""
import math

import numpy as np
import pytest

import hull as HO
from stonefish_amd import hydro as H
from stonefish_amd import meshes as M
from stonefish_amd import transforms as T

RHO = 1000.0


def _ref_meshes():
    from pathlib import Path
    return np.load(Path(__file__).resolve().parent / "golden" / "ref_meshes.npz")


def _rotated(mesh, R, t=(0.0, 0.0, 0.0)):
    v = (mesh.verts.astype(np.float64) @ np.asarray(R).T + np.asarray(t)).astype(np.float32)
    return v, mesh.faces


def _rot(rng):
    return T.quat_to_gl(T.random_unit_quaternions(rng, 1)[0], [0, 0, 0]).reshape(4, 4).T[:3, :3]


# ---- the oracle, pinned ------------------------------------------------------------

def test_hull_known_answer():
    """hull_hydro.obj: V = 0.096699 m^3 and inertia / density =
    (0.001080, 0.014772, 0.014772) (the reference test's expectations, six
    decimals)."""
    z = _ref_meshes()
    p = HO.physical_properties(z["hullhydro_v"], z["hullhydro_f"], 0.0, 1.0)
    assert abs(p["volume"] - 0.096699) < 1e-6
    np.testing.assert_allclose(sorted(p["Ipri"]), [0.001080, 0.014772, 0.014772], atol=5e-7)


def test_box_closed_forms():
    """A box 2a x 2b x 2c offset by t and rotated by R: V = 8abc, A = 8(ab + bc
    + ca), CG = t, principal moments m/3 (b^2 + c^2) etc. along R's columns."""
    a, b, c = 0.7, 0.4, 0.2
    rng = np.random.default_rng(2)
    R = _rot(rng)
    t = np.array([0.3, -1.2, 0.5])
    v, f = _rotated(M.box(a, b, c, 4), R, t)
    p = HO.physical_properties(v, f, 0.0, RHO)
    m = RHO * 8 * a * b * c
    np.testing.assert_allclose(p["volume"], 8 * a * b * c, rtol=1e-6)
    np.testing.assert_allclose(p["surface"], 8 * (a * b + b * c + c * a), rtol=1e-6)
    np.testing.assert_allclose(p["CG"], t, atol=1e-6)
    moments = np.array([m / 3 * (b * b + c * c), m / 3 * (a * a + c * c), m / 3 * (a * a + b * b)])
    np.testing.assert_allclose(sorted(p["Ipri"]), sorted(moments), rtol=1e-5)
    Irot = np.array(p["Irot"])
    np.testing.assert_allclose(Irot.T @ Irot, np.eye(3), atol=1e-9)
    for k in range(3):   # each principal axis is one of the box's axes, with its moment
        axis = Irot[:, k]
        j = int(np.argmax(np.abs(R.T @ axis)))
        assert abs(abs(R[:, j] @ axis) - 1) < 1e-6
        np.testing.assert_allclose(p["Ipri"][k], moments[j], rtol=1e-5)


def test_thin_spherical_shell():
    """A shell of thickness t << r: mass rho A t and CG at the centre.  The
    reference's moments subtract two solids whose faces are translated by
    -+t/2 along their normals (GeometryFileUtil.cpp:403-486): the polar moment
    of a tetra from the centre to a face at distance h grows as A h^3 / 5, so
    the difference is 3/5 A h^2 t per face, three fifths of a true thin
    shell's: (3/5)(2/3) m r^2 = (2/5) m r^2, which the restatement must
    reproduce (within the polyhedral error)."""
    r, th = 1.0, 1e-3
    mesh = M.ico_ellipsoid(r, r, r, 4)
    p = HO.physical_properties(mesh.verts, mesh.faces, th, RHO)
    A = p["surface"]
    np.testing.assert_allclose(p["mass"], RHO * A * th, rtol=1e-12)
    np.testing.assert_allclose(p["CG"], 0, atol=1e-9)
    np.testing.assert_allclose(p["Ipri"], 0.4 * p["mass"] * r * r, rtol=5e-3)


def test_euler_zyx_matches_axis_rotations():
    """Quaternion(yaw, pitch, roll) under BT_EULER_DEFAULT_ZYX is
    Rz(yaw) Ry(pitch) Rx(roll); the cylinder frames send z to +x and -y."""
    yz, py, rx = 0.3, -0.7, 1.1
    Rz = np.array([[math.cos(yz), -math.sin(yz), 0], [math.sin(yz), math.cos(yz), 0], [0, 0, 1]])
    Ry = np.array([[math.cos(py), 0, math.sin(py)], [0, 1, 0], [-math.sin(py), 0, math.cos(py)]])
    Rx = np.array([[1, 0, 0], [0, math.cos(rx), -math.sin(rx)], [0, math.sin(rx), math.cos(rx)]])
    np.testing.assert_allclose(HO.euler_zyx_matrix(yz, py, rx), Rz @ Ry @ Rx, atol=1e-15)
    np.testing.assert_allclose(np.array(HO.euler_zyx_matrix(0, math.pi / 2, 0)) @ [0, 0, 1], [1, 0, 0], atol=1e-15)
    np.testing.assert_allclose(np.array(HO.euler_zyx_matrix(0, 0, math.pi / 2)) @ [0, 0, 1], [0, -1, 0], atol=1e-15)


def test_sphere_and_cylinder_approx_closed_forms():
    """Sphere: radius = the farthest vertex (a box's half diagonal), Cd = 1,
    Cf = 0.1, and the added mass with the reference's squared radius.
    Cylinder along each axis: radius and length of the mesh, Cd (0.5, 0.5, 1)
    carried to the body axes."""
    a, b, c = 0.7, 0.4, 0.2
    box = M.box(a, b, c, 2)
    s = HO.fluid_dynamics_approx(box.verts, HO.SPHERE)
    r2 = a * a + b * b + c * c
    np.testing.assert_allclose(s["params"][0], math.sqrt(r2), rtol=1e-6)
    np.testing.assert_allclose(s["added_mass"], 2 * math.pi * RHO * r2 ** 3 / 3, rtol=1e-6)
    assert s["Cd"] == [1.0, 1.0, 1.0] and np.allclose(s["Cf"], 0.1)
    for axis in range(3):
        R, L = 0.3, 2.0
        cyl = M.cylinder(R, L, 64)
        perm = [[2, 0, 1], [0, 2, 1], [0, 1, 2]][axis]   # move the z axis to `axis`
        v = cyl.verts[:, perm].astype(np.float32)
        cy = HO.fluid_dynamics_approx(v, HO.CYLINDER)
        np.testing.assert_allclose(cy["params"][:2], [R, L], rtol=1e-6)
        H_ = np.array(cy["T_CG2H"]).reshape(4, 4).T[:3, :3]
        assert abs(abs((H_ @ [0, 0, 1])[axis]) - 1) < 1e-12
        cd = np.full(3, 0.5)
        cd[axis] = 1.0
        np.testing.assert_allclose(cy["Cd"], cd, atol=1e-12)


@pytest.mark.parametrize("half", [(1.0, 0.3, 0.25), (0.5, 0.5, 0.2), (0.2, 0.6, 0.4)])
def test_ellipsoid_approx_recovers_ellipsoids(half):
    """The minimum-volume axis-aligned ellipsoid of an ellipsoid mesh is the
    ellipsoid: the iteration stops inside its 20 % volume tolerance, so the
    semi-axes are within a few percent and Cd follows 1 / semi-axis."""
    mesh = M.ico_ellipsoid(*half, 3)
    e = HO.fluid_dynamics_approx(mesh.verts, HO.ELLIPSOID)
    vol = np.prod(e["params"]) / np.prod(half)
    assert 0.99 <= vol <= 1.2, (e["params"], half)
    np.testing.assert_allclose(e["params"], half, rtol=0.1)
    cd = 1 / np.array(e["params"])
    np.testing.assert_allclose(e["Cd"], cd / cd.max(), rtol=1e-12)


# ---- the library against the oracle ------------------------------------------------

def _close(a, b, rtol=1e-12, atol=1e-14):
    np.testing.assert_allclose(np.asarray(a, dtype=float), np.asarray(b, dtype=float), rtol=rtol, atol=atol)


def _meshes():
    z = _ref_meshes()
    rng = np.random.default_rng(7)
    out = [("hull_hydro", z["hullhydro_v"], z["hullhydro_f"]), ("torus", z["torusR1r025_v"], z["torusR1r025_f"]),
           ("icosphere", z["icosphere_v"], z["icosphere_f"])]
    v, f = _rotated(M.box(0.7, 0.4, 0.2, 3), _rot(rng), [0.2, 0.1, -0.3])
    out.append(("rotated box", v, f))
    e = M.ico_ellipsoid(1.0, 0.3, 0.25, 3)
    out.append(("ellipsoid", e.verts, e.faces))
    return out


@pytest.mark.parametrize("thickness", [0.0, 0.01])
def test_physical_properties_match_oracle(thickness):
    for name, v, f in _meshes():
        g = H.compute_physical_properties(v, f, thickness, RHO)
        o = HO.physical_properties(v.astype(np.float32), f, thickness, RHO)
        for k in ("mass", "volume", "surface"):
            _close(g[k], o[k], rtol=1e-12)
        scale = np.abs(o["Ipri"]).max()
        _close(g["CG"], o["CG"], atol=1e-12)
        _close(g["Ipri"], o["Ipri"], rtol=1e-9, atol=1e-12 * scale)
        _close(g["Irot"], o["Irot"], atol=1e-8)


def test_fluid_approx_matches_oracle():
    rng = np.random.default_rng(11)
    for name, v, f in _meshes():
        Tcg2c = T.quat_to_gl(T.random_unit_quaternions(rng, 1)[0], rng.uniform(-0.2, 0.2, 3))
        Tcg2o = T.quat_to_gl(T.random_unit_quaternions(rng, 1)[0], rng.uniform(-0.2, 0.2, 3))
        pcb = rng.uniform(-0.05, 0.05, 3)
        for kind in (HO.AUTO, HO.SPHERE, HO.CYLINDER, HO.ELLIPSOID):
            g = H.compute_fluid_dynamics_approx(v, kind, Tcg2c, pcb, Tcg2o, RHO)
            o = HO.fluid_dynamics_approx(v.astype(np.float32), kind, Tcg2c, pcb, Tcg2o, RHO)
            assert g["type"] == o["type"] and g["iterations"] == o["iterations"], (name, kind)
            for k in ("params", "added_mass", "added_inertia", "Cd", "Cf", "T_CG2H"):
                _close(g[k], o[k], rtol=1e-10, atol=1e-12)


def test_argument_checks():
    from stonefish_amd import _native as N
    v = np.zeros((1, 3), np.float32)
    with pytest.raises(N.SfhError):
        H.compute_fluid_dynamics_approx(v)          # fewer than two vertices
    with pytest.raises(N.SfhError):
        H.compute_fluid_dynamics_approx(np.zeros((4, 3), np.float32), kind=7)
    with pytest.raises(N.SfhError):
        H.compute_physical_properties(np.zeros((3, 3), np.float32), np.array([[0, 1, 5]]))


# ---- CorrectAerodynamicForces ---------------------------------------------------------

def test_correct_aero_closed_forms_and_oracle():
    """Sphere: unchanged.  Cylinder with the drag along its axis: factor 1,
    across it: 0.5.  Ellipsoid: |Cd . d| with Cd = (1/a, 1/b, 1/c) / max.
    The library against the restatement on random frames and vectors."""
    rng = np.random.default_rng(5)
    Rq = T.random_unit_quaternions(rng, 1)[0]
    H_ = T.quat_to_gl(Rq, [0.1, 0.2, 0.3])
    R = np.array(H_).reshape(4, 4).T[:3, :3]
    Tda0 = np.array([0.3, -0.2, 0.5])
    for kind, params, F_local, factor in ((HO.SPHERE, [0.5, 0, 0], [1.0, 2.0, 3.0], 1.0),
                                          (HO.CYLINDER, [0.3, 2.0, 0], [0, 0, -4.0], 1.0),
                                          (HO.CYLINDER, [0.3, 2.0, 0], [0, 2.5, 0], 0.5),
                                          (HO.ELLIPSOID, [1.0, 0.5, 0.25], [0, 0, 3.0], 1.0),
                                          (HO.ELLIPSOID, [1.0, 0.5, 0.25], [-2.0, 0, 0], 0.25)):
        F = R @ np.array(F_local, dtype=float)
        Fd, Td = F.copy(), Tda0.copy()
        H.CorrectAerodynamicForces(kind, params, H_, Fd, Td)
        np.testing.assert_allclose(Fd, factor * F, rtol=1e-12)
        np.testing.assert_allclose(Td, factor * Tda0, rtol=1e-12)
    for _ in range(20):
        kind = int(rng.integers(0, 4))
        params = rng.uniform(0.1, 2.0, 3)
        Hr = T.quat_to_gl(T.random_unit_quaternions(rng, 1)[0], rng.uniform(-1, 1, 3))
        F, Tq = rng.normal(size=3), rng.normal(size=3)
        if _ == 0:
            F[:] = 0.0      # safeNormalize falls back to +x
        Fo, To = HO.correct_aero(kind, params, Hr, F, Tq)
        Fd, Td = F.copy(), Tq.copy()
        H.CorrectAerodynamicForces(kind, params, Hr, Fd, Td)
        _close(Fd, Fo, rtol=1e-13)
        _close(Td, To, rtol=1e-13)
