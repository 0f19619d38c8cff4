"""The shared per-face integrand (sfh_face.h), built for the host, against the
FP64 oracle: checks the mesh-frame formulation (clip cases, tetra volume,
lever arms, moment transport) before any GPU run.  Same 1e-6 bound as the GPU
parity tests."""
import ctypes as C
import subprocess
from pathlib import Path

import numpy as np
import pytest

import oracle as O
from stonefish_amd import meshes as M
from stonefish_amd import transforms as T
from stonefish_amd import workloads as WL
from util import bound_ratio

HERE = Path(__file__).resolve().parent


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    out = tmp_path_factory.mktemp("harness") / "libharness.so"
    subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off", "-o", str(out),
                    str(HERE / "face_harness.cpp")], check=True)
    lib = C.CDLL(str(out))
    dp = C.POINTER(C.c_double)
    lib.harness_body.argtypes = [C.c_int, C.POINTER(C.c_float), C.POINTER(C.c_uint32), C.c_uint32, dp, dp, dp, dp,
                                 C.c_double, C.c_double, dp]
    return lib


def run(lib, mode, tm, Tcg, Tc, v, w):
    out = np.zeros(24)
    vv = np.ascontiguousarray(tm.verts, dtype=np.float32)
    ff = np.ascontiguousarray(tm.faces, dtype=np.uint32)
    d = lambda a: np.ascontiguousarray(a, dtype=np.float64).ctypes.data_as(C.POINTER(C.c_double))  # noqa: E731
    args = [np.ascontiguousarray(x, dtype=np.float64) for x in (Tcg, Tc, v, w)]
    lib.harness_body(mode, vv.ctypes.data_as(C.POINTER(C.c_float)), ff.ctypes.data_as(C.POINTER(C.c_uint32)),
                     ff.shape[0], *[a.ctypes.data_as(C.POINTER(C.c_double)) for a in args], 1000.0, 9.81,
                     out.ctypes.data_as(C.POINTER(C.c_double)))
    del d
    return out


def check(g, r, x, names):
    worst, plain = bound_ratio(g, r, x, names=names)
    assert max(worst.values()) <= 1.0, (worst, plain)
    return worst, plain


def test_submerged_capsule_host(harness):
    w = WL.capsule_submerged()
    tm = w.bodies[0].mesh
    om, oo = O.Mesh.from_tri(tm), O.Ocean()
    st = WL.random_states(w, 12)[:, 0, :]
    g = np.stack([run(harness, 1, tm, s[:16], s[:16], s[16:19], s[19:22]) for s in st])
    rr = [O.submerged(64, om, oo, s[:16], s[:16], s[16:19], s[19:22], full=True) for s in st]
    r = np.stack([a for a, _ in rr])
    x = np.stack([b for _, b in rr])
    r[:, 20] = 60
    check(g, r, x, ("Fdq", "Tdq", "Fdf", "Tdf"))


@pytest.mark.parametrize("offset", [(0.0, 0.0), (37.0, -512.0)])
def test_surface_ellipsoid_host(harness, offset):
    tm = M.ico_ellipsoid(1.0, 0.3, 0.25, 4)
    om, oo = O.Mesh.from_tri(tm), O.Ocean()
    rng = np.random.default_rng(4)
    g, r, x = [], [], []
    for _ in range(12):
        Tcg = T.quat_to_gl(T.random_unit_quaternions(rng, 1)[0], [offset[0], offset[1], rng.uniform(-0.2, 0.2)])
        # a C frame offset from the CG exercises the lever-arm transport
        Tc = T.gl_mul(Tcg, T.translation_gl([0.05, -0.02, 0.01]))
        v, om_ = rng.uniform(-2, 2, 3), rng.uniform(-1, 1, 3)
        g.append(run(harness, 2, tm, Tcg, Tc, v, om_))
        a, b = O.surface(64, 1, 1, om, oo, Tcg, Tc, v, om_, full=True)
        r.append(a)
        x.append(b)
    g, r, x = np.stack(g), np.stack(r), np.stack(x)
    check(g, r, x, ("Fb", "Tb", "Fdq", "Tdq", "Fdf", "Tdf"))
    np.testing.assert_allclose(g[:, 18], r[:, 18], rtol=1e-6)
    np.testing.assert_allclose(g[:, 19], r[:, 19], rtol=1e-6)


def test_fleet_crossing_host(harness):
    """Crossing bodies of a fleet (flat sea), including bodies that are fully
    under water but classified crossing by their loose AABB (zero buoyancy
    torque) and bodies with thin wet slivers."""
    w = WL.fleet(48, seed=11, f_lo=200, f_hi=4000, waves=False)
    st = WL.random_states(w, 2)[0]
    oo = O.Ocean()
    g, r, x = [], [], []
    for i, b in enumerate(w.bodies):
        om = O.Mesh.from_tri(b.mesh)
        ob = O.make_body(om, Cd=b.Cd, Cf=b.Cf, volume=b.volume, surface=b.surface)
        if O.load().orc_classify_64(C.byref(ob), C.byref(oo.s), O._d(st[i, :16])) != 2:
            continue
        g.append(run(harness, 2, b.mesh, st[i, :16], st[i, :16], st[i, 16:19], st[i, 19:22]))
        a, e = O.surface(64, 1, 1, om, oo, st[i, :16], st[i, :16], st[i, 16:19], st[i, 19:22], full=True)
        r.append(a)
        x.append(e)
    assert len(g) > 10
    check(np.stack(g), np.stack(r), np.stack(x), ("Fb", "Tb", "Fdq", "Tdq", "Fdf", "Tdf"))


def test_gelu_against_erf(harness):
    """The surrogate's GELU (erf by Abramowitz-Stegun 7.1.26, |error| <= 1.5e-7)
    against the exact erf in FP64: value and derivative within 5e-7 of the
    scale (single-precision rounding included)."""
    from math import erf, exp, pi, sqrt
    z = np.concatenate([np.linspace(-8, 8, 4001), [0.0, -0.0, 1e-6, -1e-6, 30.0, -30.0]]).astype(np.float32)
    g = np.zeros_like(z)
    dg = np.zeros_like(z)
    fp = C.POINTER(C.c_float)
    harness.harness_gelu.argtypes = [fp, C.c_uint32, fp, fp]
    harness.harness_gelu(z.ctypes.data_as(fp), z.size, g.ctypes.data_as(fp), dg.ctypes.data_as(fp))
    zd = z.astype(np.float64)
    ge = np.array([0.5 * t * (1 + erf(t / sqrt(2))) for t in zd])
    de = np.array([0.5 * (1 + erf(t / sqrt(2))) + t * exp(-0.5 * t * t) / sqrt(2 * pi) for t in zd])
    assert np.all(np.abs(g - ge) <= 5e-7 * np.maximum(1.0, np.abs(zd)))
    assert np.all(np.abs(dg - de) <= 5e-7 * np.maximum(1.0, np.abs(zd)))
    assert g[np.argmin(np.abs(zd - 30))] == np.float32(30.0) and g[np.argmin(np.abs(zd + 30))] == 0.0
