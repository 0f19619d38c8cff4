"""Pins the FP64 oracle (oracle/sfh_oracle.c) before it is trusted as the parity
target:
  - LinearMath restatement against the reference's own Bullet LinearMath,
    compiled from the reference checkout into oracle/_ref (skipped where that
    library is absent);
  - the known answer the reference's FluidDynamicsTest holds for
    hull_hydro.obj (volume 0.096699 m^3, Tests/FluidDynamicsTest/
    FluidDynamicsTestManager.cpp:207-215);
  - closed forms (Archimedes, a half-submerged box, drag on a translating box,
    the correction factor);
  - invariants (horizontal translation, rotation about the vertical, face
    order, winding, mirror symmetry, pressure integration = volume buoyancy);
  - every clip case against an independent polygon clipper.
"""
import ctypes as C
from pathlib import Path

import numpy as np
import pytest

import oracle as O
from stonefish_amd import meshes as M
from stonefish_amd import transforms as T

HERE = Path(__file__).resolve().parent
REF_LIB = HERE.parent / "oracle" / "_ref" / "libbtref.so"
RHO, G = 1000.0, 9.81


def d(a):
    return O._d(a)


@pytest.fixture(scope="module")
def bt():
    if not REF_LIB.exists():
        pytest.skip("oracle/_ref (reference Bullet LinearMath) not built")
    lib = C.CDLL(str(REF_LIB))
    dp = C.POINTER(C.c_double)
    lib.btref_transform_aabb.argtypes = [dp, dp, C.c_double, dp, dp, dp]
    return lib


def arr(n):
    return (C.c_double * n)()


# ---------------------------------------------------------------- LinearMath

def test_quat_to_gl_matches_bullet(bt):
    lib = O.load()
    rng = np.random.default_rng(0)
    for _ in range(200):
        q = rng.normal(size=4) * rng.uniform(0.1, 3.0)   # not normalised: setRotation divides by |q|^2
        o = rng.uniform(-100, 100, 3)
        a, b = arr(16), arr(16)
        lib.orc_quat_to_gl(d(q), d(o), a)
        bt.btref_quat_to_gl(d(q), d(o), b)
        np.testing.assert_array_equal(np.array(a), np.array(b))
        # the product's input former agrees too
        np.testing.assert_allclose(T.quat_to_gl(q, o), np.array(b), rtol=0, atol=1e-14)


def test_transform_ops_match_bullet(bt):
    lib = O.load()
    rng = np.random.default_rng(1)
    for _ in range(200):
        A = T.quat_to_gl(rng.normal(size=4), rng.uniform(-10, 10, 3))
        B = T.quat_to_gl(rng.normal(size=4), rng.uniform(-10, 10, 3))
        x, y = arr(16), arr(16)
        lib.orc_gl_mul(d(A), d(B), x)
        bt.btref_gl_mul(d(A), d(B), y)
        np.testing.assert_allclose(np.array(x), np.array(y), rtol=0, atol=1e-13)
        lib.orc_gl_inverse(d(A), x)
        bt.btref_gl_inverse(d(A), y)
        np.testing.assert_allclose(np.array(x), np.array(y), rtol=0, atol=1e-13)
        i1, i2 = arr(9), arr(9)
        lib.orc_basis_inverse(d(A), i1)
        bt.btref_basis_inverse(d(A), i2)
        np.testing.assert_allclose(np.array(i1), np.array(i2), rtol=0, atol=1e-14)
        mn, mx = rng.uniform(-2, 0, 3), rng.uniform(0, 2, 3)
        m = rng.uniform(0, 0.01)
        a1, b1, a2, b2 = arr(3), arr(3), arr(3), arr(3)
        lib.orc_transform_aabb(d(mn), d(mx), m, d(A), a1, b1)
        bt.btref_transform_aabb(d(mn), d(mx), m, d(A), a2, b2)
        np.testing.assert_allclose(np.array(a1), np.array(a2), rtol=0, atol=1e-13)
        np.testing.assert_allclose(np.array(b1), np.array(b2), rtol=0, atol=1e-13)


def test_safe_normalize_matches_bullet(bt):
    lib = O.load()
    rng = np.random.default_rng(2)
    cases = [np.zeros(3), np.array([1e-170, 0, 0]), np.array([3e-9, -2e-9, 1e-9])] + \
        [rng.normal(size=3) * 10 ** rng.uniform(-8, 8) for _ in range(50)]
    for v in cases:
        a, b = d(v), d(v)
        lib.orc_safe_normalize(a)
        bt.btref_safe_normalize(b)
        np.testing.assert_array_equal(np.array(a), np.array(b))


def test_correct_matches_bullet_pieces(bt):
    """CorrectHydrodynamicForces (SolidEntity.cpp:1263-1287) assembled from the
    reference's own inverse, basis product and safeNormalize."""
    rng = np.random.default_rng(3)
    oc = O.Ocean()
    for _ in range(50):
        TO = T.quat_to_gl(rng.normal(size=4), rng.uniform(-5, 5, 3))
        Cd, Cf = rng.uniform(0.05, 1, 3), rng.uniform(0.005, 0.1, 3)
        vecs = [rng.normal(size=3) * 10 ** rng.uniform(-3, 3) for _ in range(4)]
        if rng.random() < 0.2:
            vecs[1] = np.zeros(3)   # safeNormalize of zero -> (1, 0, 0)
        got = O.correct(oc, *vecs, Cd, Cf, TO)
        inv = arr(9)
        bt.btref_basis_inverse(d(TO), inv)
        for q in range(4):
            x = vecs[q]
            y = arr(3)
            bt.btref_basis_mul_vec(inv, d(x), y)
            nv = d(np.array(y))
            bt.btref_safe_normalize(nv)
            C_ = Cd if q < 2 else Cf
            c = float(np.abs(np.array(nv)) @ C_)
            k = 0.5 * RHO * c if q < 2 else RHO * c
            np.testing.assert_allclose(got[q], k * x, rtol=1e-15, atol=0)


def test_correct_closed_form():
    oc = O.Ocean()
    Cd, Cf = np.array([0.3, 0.6, 0.9]), np.array([0.03, 0.06, 0.09])
    I = T.identity_gl()
    F = O.correct(oc, [2.0, 0, 0], [0, 3.0, 0], [0, 0, -1.0], [0, 0, 0], Cd, Cf, I)
    np.testing.assert_allclose(F[0], [0.5 * RHO * 0.3 * 2.0, 0, 0])
    np.testing.assert_allclose(F[1], [0, 0.5 * RHO * 0.6 * 3.0, 0])
    np.testing.assert_allclose(F[2], [0, 0, -RHO * 0.09])
    np.testing.assert_allclose(F[3], [0, 0, 0])


# ---------------------------------------------------------------- known answers

def ref_meshes():
    return np.load(HERE / "golden" / "ref_meshes.npz")


def test_hull_volume_known_answer():
    """hull_hydro.obj: V = 0.096699 m^3 with the reference test's tolerance
    (testScalar, absolute 1e-6), from the surface loop of a fully submerged
    hull in random poses."""
    z = ref_meshes()
    m = O.Mesh(z["hullhydro_v"], z["hullhydro_f"])
    oc = O.Ocean()
    rng = np.random.default_rng(4)
    for _ in range(5):
        Tcg = T.quat_to_gl(T.random_unit_quaternions(rng, 1)[0], [rng.uniform(-50, 50), rng.uniform(-50, 50), 10.0])
        r = O.surface(64, 1, 1, m, oc, Tcg, Tcg, [0, 0, 0], [0, 0, 0])
        assert abs(r[19] - 0.096699) < 1e-6, r[19]
        np.testing.assert_allclose(r[0:3], [0, 0, -RHO * G * r[19]], rtol=1e-12, atol=1e-9)


# ---------------------------------------------------------------- closed forms

def test_archimedes_box_offset():
    """Fully submerged box whose centre is offset by a from the CG: Fb = -rho g V,
    Tb = (a x) x Fb (centre of buoyancy at the box centre)."""
    hx, hy, hz = (float(np.float32(x)) for x in (0.6, 0.4, 0.3))   # the mesh holds float vertices
    tm = M.box(hx, hy, hz, 3)
    m = O.Mesh.from_tri(tm)
    oc = O.Ocean()
    a = np.array([0.25, -0.1, 0.05])
    Tcg = T.translation_gl([3.0, 4.0, 7.0])
    Tc = T.gl_mul(Tcg, T.translation_gl(a))
    r = O.surface(64, 1, 1, m, oc, Tcg, Tc, [0, 0, 0], [0, 0, 0])
    V = 8 * hx * hy * hz
    np.testing.assert_allclose(r[19], V, rtol=1e-12)
    Fb = np.array([0, 0, -RHO * G * V])
    np.testing.assert_allclose(r[0:3], Fb, rtol=1e-12)
    np.testing.assert_allclose(r[3:6], np.cross(a, Fb), rtol=1e-9, atol=1e-9)


@pytest.mark.parametrize("depth", [-0.2, 0.0, 0.17])
def test_half_submerged_box(depth):
    """Axis-aligned box, centre at depth `depth` (NED), offset a along x from
    the CG: wet prism V = 4 hx hy (hz + depth), centre of buoyancy at
    z = (hz + depth)/2, Tb = (a, 0, zc - zcg) x Fb."""
    hx, hy, hz = (float(np.float32(x)) for x in (0.5, 0.3, 0.25))
    tm = M.box(hx, hy, hz, 4)
    m = O.Mesh.from_tri(tm)
    oc = O.Ocean()
    a = 0.2
    Tcg = T.translation_gl([10.0 - a, -3.0, depth])
    Tc = T.translation_gl([10.0, -3.0, depth])
    r = O.surface(64, 1, 1, m, oc, Tcg, Tc, [0, 0, 0], [0, 0, 0])
    V = 4 * hx * hy * (hz + depth)
    np.testing.assert_allclose(r[19], V, rtol=1e-12)
    Fb = np.array([0, 0, -RHO * G * V])
    np.testing.assert_allclose(r[0:3], Fb, rtol=1e-12)
    arm = np.array([a, 0.0, (hz + depth) / 2.0 - depth])
    np.testing.assert_allclose(r[3:6], np.cross(arm, Fb), rtol=1e-9, atol=1e-8)
    # wetted area: bottom + the wet strips of the four sides
    Swet = 4 * hx * hy + 2 * (2 * hx + 2 * hy) * (hz + depth)
    np.testing.assert_allclose(r[18], Swet, rtol=1e-12)


def test_translating_box_drag():
    """Box translating along +x at u, no rotation, fully submerged: only the
    front face approaches the fluid, quadratic term = vc |vc| (-vc.n) A summed
    = (-u^3 A_front, 0, 0); skin friction from the four side faces =
    (-u A_sides, 0, 0); no torque about the centre."""
    hx, hy, hz = (float(np.float32(x)) for x in (0.7, 0.4, 0.2))
    m = O.Mesh.from_tri(M.box(hx, hy, hz, 5))
    oc = O.Ocean()
    u = 1.7
    Tcg = T.translation_gl([0, 0, 5.0])
    r = O.submerged(64, m, oc, Tcg, Tcg, [u, 0, 0], [0, 0, 0])
    np.testing.assert_allclose(r[6:9], [-u ** 3 * 4 * hy * hz, 0, 0], rtol=1e-12, atol=1e-10)
    np.testing.assert_allclose(r[12:15], [-u * (8 * hx * hz + 8 * hx * hy), 0, 0], rtol=1e-12, atol=1e-10)
    np.testing.assert_allclose(r[9:12], 0, atol=1e-9)
    np.testing.assert_allclose(r[15:18], 0, atol=1e-9)


# ---------------------------------------------------------------- invariants

def random_pose(rng, depth=5.0):
    return T.quat_to_gl(T.random_unit_quaternions(rng, 1)[0], [rng.uniform(-3, 3), rng.uniform(-3, 3), depth])


def test_horizontal_translation_invariance():
    tm = M.ico_ellipsoid(1.0, 0.4, 0.3, 3)
    m = O.Mesh.from_tri(tm)
    oc = O.Ocean()
    rng = np.random.default_rng(5)
    for _ in range(5):
        Tcg = random_pose(rng, depth=rng.uniform(-0.2, 0.2))
        v, w = rng.uniform(-2, 2, 3), rng.uniform(-1, 1, 3)
        r1 = O.surface(64, 1, 1, m, oc, Tcg, Tcg, v, w)
        T2 = Tcg.copy()
        T2[12:14] += [123.0, -77.0]
        r2 = O.surface(64, 1, 1, m, oc, T2, T2, v, w)
        np.testing.assert_allclose(r2[:20], r1[:20], rtol=1e-10, atol=1e-8)


def test_vertical_axis_rotation_covariance():
    """Rotating the body, its velocities and the frame about the vertical
    leaves the flat sea unchanged, so forces and torques rotate."""
    tm = M.ico_ellipsoid(1.0, 0.4, 0.3, 3)
    m = O.Mesh.from_tri(tm)
    oc = O.Ocean()
    rng = np.random.default_rng(6)
    for _ in range(5):
        Tcg = random_pose(rng, depth=rng.uniform(-0.2, 0.2))
        Tcg[12:14] = 0.0
        v, w = rng.uniform(-2, 2, 3), rng.uniform(-1, 1, 3)
        th = rng.uniform(0, 2 * np.pi)
        Q = T.quat_to_gl([0, 0, np.sin(th / 2), np.cos(th / 2)], [0, 0, 0])
        R = T.basis(Q)
        T2 = T.gl_mul(Q, Tcg)
        r1 = O.surface(64, 1, 1, m, oc, Tcg, Tcg, v, w)
        r2 = O.surface(64, 1, 1, m, oc, T2, T2, R @ v, R @ w)
        for k in range(6):
            np.testing.assert_allclose(r2[3 * k:3 * k + 3], R @ r1[3 * k:3 * k + 3], rtol=1e-9,
                                       atol=1e-9 * max(1.0, np.abs(r1[3 * k:3 * k + 3]).max()))


def test_submerged_rotation_covariance():
    """Fully submerged: the drag loop is invariant under any rotation."""
    m = O.Mesh.from_tri(M.capsule(0.15, 1.6))
    oc = O.Ocean()
    rng = np.random.default_rng(7)
    for _ in range(5):
        Tcg = random_pose(rng)
        Tcg[12:15] = 0.0
        v, w = rng.uniform(-2, 2, 3), rng.uniform(-1, 1, 3)
        Q = T.quat_to_gl(T.random_unit_quaternions(rng, 1)[0], [0, 0, 0])
        R = T.basis(Q)
        T2 = T.gl_mul(Q, Tcg)
        r1 = O.submerged(64, m, oc, Tcg, Tcg, v, w)
        r2 = O.submerged(64, m, oc, T2, T2, R @ v, R @ w)
        for k in range(2, 6):
            np.testing.assert_allclose(r2[3 * k:3 * k + 3], R @ r1[3 * k:3 * k + 3], rtol=1e-9,
                                       atol=1e-12 * max(1.0, np.abs(r1).max()))


def test_face_order_and_winding():
    tm = M.ico_ellipsoid(1.0, 0.4, 0.3, 3)
    oc = O.Ocean()
    rng = np.random.default_rng(8)
    Tcg = random_pose(rng, depth=0.1)
    v, w = rng.uniform(-2, 2, 3), rng.uniform(-1, 1, 3)
    r1 = O.surface(64, 1, 1, O.Mesh.from_tri(tm), oc, Tcg, Tcg, v, w)
    perm = rng.permutation(tm.nf)
    r2 = O.surface(64, 1, 1, O.Mesh(tm.verts, tm.faces[perm]), oc, Tcg, Tcg, v, w)
    np.testing.assert_allclose(r2[:20], r1[:20], rtol=1e-11, atol=1e-9)
    # reversed winding: the volume integral turns negative, so buoyancy and
    # Vsub are left untouched (SolidEntity.cpp:1690); wetted area is the same
    init = np.full(24, 7.0)
    r3 = O.surface(64, 1, 1, O.Mesh(tm.verts, tm.faces[:, ::-1].copy()), oc, Tcg, Tcg, v, w, init=init)
    assert np.all(r3[0:6] == 7.0) and r3[19] == 7.0
    np.testing.assert_allclose(r3[18], r1[18], rtol=1e-12)


def test_mirror_symmetry_zero_components():
    """An ellipsoid moving in its x-z symmetry plane with rotation about y: the
    lateral force and the x, z torques vanish."""
    m = O.Mesh.from_tri(M.uv_ellipsoid(1.0, 0.4, 0.3, 64, 33))
    oc = O.Ocean()
    Tcg = T.translation_gl([0, 0, 0.1])
    r = O.surface(64, 1, 1, m, oc, Tcg, Tcg, [1.2, 0, 0.3], [0, 0.4, 0])
    # (the quads of the mesh are split along one diagonal, which the mirror
    # maps to the other: the discrete body is symmetric to about 1e-9)
    scale = np.abs(r[:18]).max()
    for k in range(6):
        if k % 2 == 0:
            assert abs(r[3 * k + 1]) < 1e-7 * scale   # force y
        else:
            assert abs(r[3 * k]) < 1e-7 * scale and abs(r[3 * k + 2]) < 1e-7 * scale


def test_pressure_integration_equals_volume_buoyancy():
    """With waves on but a constant elevation, pressure integration over a
    fully submerged closed mesh gives the volume buoyancy exactly (centroid
    rule is exact for the linear pressure field)."""
    tm = M.ico_ellipsoid(0.8, 0.5, 0.3, 3)
    m = O.Mesh.from_tri(tm)
    h0 = 0.37
    flat = O.Ocean()
    waves = O.Ocean(wave_data=np.full((8, 8), h0, dtype=np.float32), wave_n=8, channels=1, lengths=[50.0])
    rng = np.random.default_rng(9)
    Tcg = random_pose(rng, depth=3.0)
    a = O.surface(64, 0, 1, m, flat, Tcg, Tcg, [0, 0, 0], [0, 0, 0])
    b = O.surface(64, 0, 1, m, waves, Tcg, Tcg, [0, 0, 0], [0, 0, 0])
    np.testing.assert_allclose(b[0:3], a[0:3], rtol=1e-10, atol=1e-8)
    np.testing.assert_allclose(b[3:6], a[3:6], rtol=1e-8, atol=1e-7)


# ---------------------------------------------------------------- wave query

def bilinear_periodic(grid, x, y, L):
    """Independent periodic bilinear interpolation with texel centres at
    (i + 0.5) L / n."""
    n = grid.shape[0]
    u = x / L * n - 0.5
    v = y / L * n - 0.5
    i0, j0 = int(np.floor(u)), int(np.floor(v))
    a, b = u - i0, v - j0
    g = lambda j, i: float(grid[j % n, i % n])  # noqa: E731
    return (1 - a) * (1 - b) * g(j0, i0) + a * (1 - b) * g(j0, i0 + 1) + (1 - a) * b * g(j0 + 1, i0) + \
        a * b * g(j0 + 1, i0 + 1)


def test_wave_height_query():
    rng = np.random.default_rng(10)
    n, L = 16, 40.0
    grid = rng.normal(size=(n, n)).astype(np.float32)
    oc = O.Ocean(wave_data=grid, wave_n=n, channels=1, lengths=[L])
    lib = O.load()
    pts = list(rng.uniform(-200, 200, (200, 2))) + [((i + 0.5) * L / n, (j + 0.5) * L / n) for i, j in
                                                    [(0, 0), (3, 7), (15, 15)]]
    for x, y in pts:
        h = lib.orc_wave_height_64(C.byref(oc.s), x, y)
        np.testing.assert_allclose(h, -bilinear_periodic(grid, x, y, L), rtol=1e-12, atol=1e-12)
    # texel centres reproduce the texel values
    assert abs(lib.orc_wave_height_64(C.byref(oc.s), 3.5 * L / n, 7.5 * L / n) + grid[7, 3]) < 1e-12
    # two layers sum with their own sizes
    g2 = rng.normal(size=(n, n, 4)).astype(np.float32)
    oc2 = O.Ocean(wave_data=g2, wave_n=n, channels=4, lengths=[L, 7.0])
    for x, y in rng.uniform(-50, 50, (50, 2)):
        h = lib.orc_wave_height_64(C.byref(oc2.s), x, y)
        e = -bilinear_periodic(g2[:, :, 0], x, y, L) - bilinear_periodic(g2[:, :, 1], x, y, 7.0)
        np.testing.assert_allclose(h, e, rtol=1e-12, atol=1e-12)


# ---------------------------------------------------------------- clip cases

def clip_polygon(P, depth):
    """Sutherland-Hodgman clip of polygon P against depth >= 0."""
    out = []
    n = len(P)
    for k in range(n):
        a, b = P[k], P[(k + 1) % n]
        da, db = depth(a), depth(b)
        if da >= 0:
            out.append(a)
        if (da >= 0) != (db >= 0):
            t = da / (da - db)
            out.append(a + t * (b - a))
    return out


def poly_props(P):
    P = [np.asarray(p, dtype=np.float64) for p in P]
    area_vec = np.zeros(3)
    for k in range(1, len(P) - 1):
        area_vec += 0.5 * np.cross(P[k] - P[0], P[k + 1] - P[0])
    return area_vec


CASES = {  # depths of (p1, p2, p3): which clip case of SolidEntity.cpp:1362-1647
    "A": (-0.3, -0.2, 0.4), "B": (-0.3, 0.5, -0.1), "C": (-0.4, 0.3, 0.2), "D": (0.3, -0.2, -0.5),
    "E": (0.3, -0.2, 0.25), "F": (0.2, 0.35, -0.3), "G": (0.1, 0.2, 0.3), "zero_is_wet": (0.0, -0.2, -0.3),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_clip_cases_against_polygon_clipper(case):
    """One tilted triangle per case: wetted area and submerged volume of the
    oracle against Sutherland-Hodgman clipping and the pyramid volume from the
    apex p0 = (x_CG, y_CG, 0)."""
    d0, d1, d2 = CASES[case]
    base = np.array([[0.0, 0.0, 0.0], [1.0, 0.1, 0.0], [0.2, 0.9, 0.0]])
    P = base.copy()
    P[:, 2] = [d0, d1, d2]
    verts = P.astype(np.float32)
    P = verts.astype(np.float64)   # the oracle sees the float vertices
    m = O.Mesh(verts, np.array([[0, 1, 2]], dtype=np.uint32))
    oc = O.Ocean()
    cg = np.array([0.3, 0.2, 0.05])
    Tcg = T.translation_gl(cg)
    Tc = T.identity_gl()
    r = O.surface(64, 0, 1, m, oc, Tcg, Tc, [0, 0, 0], [0, 0, 0], init=np.zeros(24))
    poly = clip_polygon(list(P), lambda p: p[2])
    av = poly_props(poly)
    np.testing.assert_allclose(r[18], np.linalg.norm(av), rtol=1e-12)
    p0 = np.array([cg[0], cg[1], 0.0])
    vol = np.dot(poly[0] - p0, av) / 3.0      # signed pyramid volume
    if vol * 6 > 1e-9:
        np.testing.assert_allclose(r[19], vol, rtol=1e-12)
    else:
        assert r[19] == 0.0


def test_degenerate_first_triangle_quirk():
    """Case C with the clip point on p1-p2 landing on p2 (d1 = 0): the quad's
    first triangle (p1', p2, p4) is degenerate, so the reference skips the
    face (no wetted area, no drag) after accumulating its tetra volume.  A
    second, fully wet face keeps the total volume positive; the difference with
    and without the quirky face is that face's clipped pyramid volume."""
    P = np.array([[0.0, 0.0, -0.4], [1.0, 0.1, 0.0], [0.2, 0.9, 0.3],
                  [-2.0, -2.0, 1.0], [-2.0, 2.0, 1.0], [2.0, 0.0, 1.0]], dtype=np.float32)
    F_both = np.array([[0, 1, 2], [3, 5, 4]], dtype=np.uint32)
    F_one = F_both[1:].copy()
    oc = O.Ocean()
    cg = np.array([0.3, 0.2, -1.0])
    Tcg = T.translation_gl(cg)
    a = O.surface(64, 1, 1, O.Mesh(P, F_both), oc, Tcg, T.identity_gl(), [1, 0, 0], [0, 0, 0])
    b = O.surface(64, 1, 1, O.Mesh(P, F_one), oc, Tcg, T.identity_gl(), [1, 0, 0], [0, 0, 0])
    assert b[19] > 0
    np.testing.assert_array_equal(a[18], b[18])           # skipped: no extra wetted area
    np.testing.assert_array_equal(a[6:18], b[6:18])       # and no extra drag
    quad = clip_polygon(list(P[:3].astype(np.float64)), lambda p: p[2])
    p0 = np.array([cg[0], cg[1], 0.0])
    vol = np.dot(quad[0] - p0, poly_props(quad)) / 3.0
    assert abs(vol) > 1e-3
    np.testing.assert_allclose(a[19] - b[19], vol, rtol=1e-10)   # its volume was accumulated


def test_ref32_close_to_ref64():
    """The reference's own float path stays within float error of the FP64
    restatement (a sanity check of the float flavour)."""
    m = O.Mesh.from_tri(M.ico_ellipsoid(1.0, 0.3, 0.25, 4))
    oc = O.Ocean()
    rng = np.random.default_rng(11)
    for _ in range(5):
        Tcg = random_pose(rng, depth=rng.uniform(-0.2, 0.2))
        v, w = rng.uniform(-2, 2, 3), rng.uniform(-1, 1, 3)
        a = O.surface(64, 1, 1, m, oc, Tcg, Tcg, v, w)
        b = O.surface(32, 1, 1, m, oc, Tcg, Tcg, v, w)
        for k in (0, 2, 4):
            np.testing.assert_allclose(b[3 * k:3 * k + 3], a[3 * k:3 * k + 3], rtol=1e-3,
                                       atol=1e-3 * np.abs(a[3 * k:3 * k + 3]).max())


# ---------------------------------------------------------------- velocity fields

def _velocity(ocn, p):
    out = arr(3)
    O.load().orc_fluid_velocity(C.byref(ocn.s), d(p), out)
    return np.array(out[:])


def test_jet_and_pipe_velocity_closed_forms():
    """Jet (Jet.cpp:54-75): on the axis u = 10 r / (t + 5 r) vout n (vout at
    t = 5 r); off the axis the Gaussian exp(-50 d^2 / t^2); nothing behind the
    outlet or outside the cone of radius (t + 5 r) / 5.  Pipe (Pipe.cpp:57-78):
    on the axis r1 / r(t) vin n with r(t) linear from r1 to r2; at half the
    radius the factor 0.5^gamma; nothing past either end or outside r(t).
    Fields add (Ocean.cpp:183-196) and are ignored when disabled."""
    n = np.array([0.0, 0.6, 0.8])
    c = np.array([1.0, -2.0, 3.0])
    r, vout = 0.4, 2.5
    e = np.cross(n, [1.0, 0.0, 0.0])
    e /= np.linalg.norm(e)
    jet = (1, 1, list(c) + list(n) + [r, vout])
    ocn = O.Ocean(currents=[jet], currents_enabled=True)
    np.testing.assert_allclose(_velocity(ocn, c + 5 * r * n), vout * n, rtol=1e-14)
    t = 0.7
    np.testing.assert_allclose(_velocity(ocn, c + t * n), 10 * r / (t + 5 * r) * vout * n, rtol=1e-14)
    dd = 0.05
    np.testing.assert_allclose(_velocity(ocn, c + t * n + dd * e),
                               np.exp(-50 * dd * dd / (t * t)) * 10 * r / (t + 5 * r) * vout * n, rtol=1e-12)
    assert np.all(_velocity(ocn, c - 0.1 * n) == 0)
    assert np.all(_velocity(ocn, c + t * n + 1.001 * (t + 5 * r) / 5 * e) == 0)
    p1, ln, r1, r2, vin, gam = np.array([0.0, 0.0, 0.0]), 4.0, 1.0, 0.5, 1.5, 1 / 7
    a = np.array([1.0, 0.0, 0.0])
    pipe = (2, 1, list(p1) + list(a) + [ln, r1, r2, vin, gam])
    ocn = O.Ocean(currents=[pipe], currents_enabled=True)
    rt = r1 + (r2 - r1) * 0.5
    np.testing.assert_allclose(_velocity(ocn, p1 + 2.0 * a), r1 / rt * vin * a, rtol=1e-14)
    side = np.array([0.0, 0.0, 1.0])
    np.testing.assert_allclose(_velocity(ocn, p1 + 2.0 * a + 0.5 * rt * side), 0.5 ** gam * r1 / rt * vin * a,
                               rtol=1e-12)
    assert np.all(_velocity(ocn, p1 + 4.01 * a) == 0) and np.all(_velocity(ocn, p1 - 0.01 * a) == 0)
    assert np.all(_velocity(ocn, p1 + 2.0 * a + 1.01 * rt * side) == 0)
    both = O.Ocean(currents=[pipe, (0, 1, [0.1, 0.2, 0.3]), jet], currents_enabled=True)
    q = p1 + 2.0 * a
    np.testing.assert_allclose(_velocity(both, q), r1 / rt * vin * a + [0.1, 0.2, 0.3], rtol=1e-14)
    assert np.all(_velocity(O.Ocean(currents=[pipe], currents_enabled=False), q) == 0)
    assert np.all(_velocity(O.Ocean(currents=[(2, 0, pipe[2])], currents_enabled=True), q) == 0)


def test_body_in_uniform_pipe_core_equals_moving_body():
    """A pipe with r1 = r2 and gamma = 0 is uniform (vin n) inside its radius:
    a submerged box at rest inside it feels the drag of the same box moving
    at -vin n through still water (the per-face velocity path against the
    relative-velocity closed form)."""
    hx, hy, hz = (float(np.float32(x)) for x in (0.7, 0.4, 0.2))
    m = O.Mesh.from_tri(M.box(hx, hy, hz, 5))
    rng = np.random.default_rng(3)
    Tcg = T.quat_to_gl(T.random_unit_quaternions(rng, 1)[0], [0.0, 0.0, 5.0])
    vin = 1.3
    a = np.array([0.6, 0.0, 0.8])
    pipe = (2, 1, list(np.array([0.0, 0.0, 5.0]) - 10 * a) + list(a) + [20.0, 5.0, 5.0, vin, 0.0])
    r_pipe = O.submerged(64, m, O.Ocean(currents=[pipe], currents_enabled=True), Tcg, Tcg, [0, 0, 0], [0, 0, 0])
    r_move = O.submerged(64, m, O.Ocean(), Tcg, Tcg, -vin * a, [0, 0, 0])
    np.testing.assert_allclose(r_pipe[6:18], r_move[6:18], rtol=1e-12, atol=1e-12)
    assert np.linalg.norm(r_move[6:9]) > 0.1


# ---------------------------------------------------------------- aerodynamics

def test_aero_translating_box_closed_form():
    """A box of half extents (hx, hy, hz) in any orientation translating at u
    through still air: only the upwind faces (n.u_hat > 0) act.  The
    reference's term is `vn * vn.length() * A` (SolidEntity.cpp:1913) with
    vn = (vc.n) n, vc = -u, on a glm::vec3, whose member length() is the
    component count 3 (GLSL semantics; the norm is the free glm::length), so
    F = 0.5 rho sum 3 A_i (vc.n_i) n_i = -1.5 rho |u| sum A_i (n_i.u_hat) n_i:
    linear in the speed, not quadratic, and no torque about the centre
    (SolidEntity.cpp:1865-1918).  The quadratic law -0.5 rho |u|^2 sum
    A_i (n_i.u_hat)^2 n_i (the norm read of length()) is checked to differ."""
    hx, hy, hz = (float(np.float32(x)) for x in (0.7, 0.4, 0.2))
    m = O.Mesh.from_tri(M.box(hx, hy, hz, 4))
    rng = np.random.default_rng(12)
    rho = 1.2
    for _ in range(4):
        Tcg = random_pose(rng, depth=-30.0)
        R = np.array(Tcg).reshape(4, 4).T[:3, :3]
        u = rng.uniform(-3, 3, 3)
        F, T_ = O.aero(m, rho, [], Tcg, Tcg, u, [0, 0, 0])
        U = np.linalg.norm(u)
        uh = u / U
        areas = np.array([4 * hy * hz, 4 * hx * hz, 4 * hx * hy])
        expect = np.zeros(3)
        quad = np.zeros(3)
        for k in range(3):
            for s in (1.0, -1.0):
                n = s * R[:, k]
                c = n @ uh
                if c > 0:
                    expect += -1.5 * rho * U * areas[k] * c * n
                    quad += -0.5 * rho * U * U * areas[k] * c * c * n
        np.testing.assert_allclose(F, expect, rtol=1e-9, atol=1e-12)
        np.testing.assert_allclose(T_, 0, atol=1e-9)
        assert np.linalg.norm(F - quad) > 1e-3 * np.linalg.norm(F)
        # linear law: twice the speed, twice the force
        F2, _ = O.aero(m, rho, [], Tcg, Tcg, 2.0 * u, [0, 0, 0])
        np.testing.assert_allclose(F2, 2.0 * F, rtol=1e-12, atol=1e-12)


def test_aero_face_term_is_three_area_normal_velocity():
    """One triangle in the x-y plane (unit normal +z, area 0.5) moving down
    at speed s (vc = +s z, vc.n = s > 0: not upwind, no force) and up at
    speed s (vc.n = -s < 0): q = 3 A vn = (0, 0, -1.5 s), Fda = 0.5 rho q.
    Pins the factor 3 of glm's vec3::length() and its sign."""
    m = O.Mesh(np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32), np.array([[0, 1, 2]], np.uint32))
    T = [1.0, 0, 0, 0, 0, 1.0, 0, 0, 0, 0, 1.0, 0, 0, 0, -10.0, 1.0]
    rho = 1.2
    for s in (0.5, 2.0):
        F, _ = O.aero(m, rho, [], T, T, [0, 0, -s], [0, 0, 0])
        np.testing.assert_allclose(F, 0.0, atol=0)
        F, _ = O.aero(m, rho, [], T, T, [0, 0, s], [0, 0, 0])
        np.testing.assert_allclose(F, [0.0, 0.0, -0.5 * rho * 3.0 * 0.5 * s], rtol=1e-12, atol=1e-15)


def test_aero_uniform_wind_equals_moving_body():
    """A body at rest in a uniform wind w feels what the body moving at -w in
    still air feels; a spinning sphere about its centre feels ~nothing."""
    m = O.Mesh.from_tri(M.ico_ellipsoid(1.0, 0.3, 0.25, 3))
    rng = np.random.default_rng(2)
    Tcg = random_pose(rng, depth=-10.0)
    w = np.array([1.5, -0.5, 0.2])
    F1, T1 = O.aero(m, 1.2, [(0, list(w))], Tcg, Tcg, [0, 0, 0], [0, 0, 0])
    F2, T2 = O.aero(m, 1.2, [], Tcg, Tcg, -w, [0, 0, 0])
    np.testing.assert_allclose(F1, F2, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(T1, T2, rtol=1e-12, atol=1e-12)
    s = O.Mesh.from_tri(M.ico_ellipsoid(1.0, 1.0, 1.0, 4))
    Fs, Ts = O.aero(s, 1.2, [], Tcg, Tcg, [0, 0, 0], [0.7, -0.3, 0.4])
    Fr, _ = O.aero(s, 1.2, [], Tcg, Tcg, [0.7, 0, 0], [0, 0, 0])
    assert np.linalg.norm(Fs) < 1e-3 * np.linalg.norm(Fr)
