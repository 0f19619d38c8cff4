"""ctypes binding of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check and bench.py's
CPU-baseline leg, as the checker or the timed CPU baseline.  Never used by the
product (stonefish_amd/), which runs on the GPU only.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
dp = C.POINTER(C.c_double)

W_FB, W_TB, W_FDQ, W_TDQ, W_FDF, W_TDF, W_SWET, W_VSUB = 1, 2, 4, 8, 16, 32, 64, 128


class orc_current(C.Structure):
    _fields_ = [("type", C.c_int32), ("enabled", C.c_int32), ("p", C.c_double * 12)]


class orc_ocean(C.Structure):
    _fields_ = [("density", C.c_double), ("gravity", C.c_double * 3), ("has_waves", C.c_int32),
                ("currents_enabled", C.c_int32), ("wave_data", C.POINTER(C.c_float)), ("wave_n", C.c_uint32),
                ("wave_channels", C.c_uint32), ("wave_layers", C.c_uint32), ("ncurrents", C.c_uint32),
                ("wave_len", C.c_float * 4), ("currents", C.POINTER(orc_current))]


class orc_mesh(C.Structure):
    _fields_ = [("verts", C.POINTER(C.c_float)), ("faces", C.POINTER(C.c_uint32)), ("nv", C.c_uint32),
                ("nf", C.c_uint32)]


class orc_result(C.Structure):
    _fields_ = [("Fb", C.c_double * 3), ("Tb", C.c_double * 3), ("Fdq", C.c_double * 3), ("Tdq", C.c_double * 3),
                ("Fdf", C.c_double * 3), ("Tdf", C.c_double * 3), ("Swet", C.c_double), ("Vsub", C.c_double),
                ("written", C.c_uint32), ("position", C.c_int32), ("S", C.c_double * 6), ("raw", C.c_double * 12),
                ("k", C.c_double * 4)]

    def extras(self) -> np.ndarray:
        """22 doubles: condition scales S[6], raw drag[12], Correct factors k[4]."""
        return np.concatenate([list(self.S), list(self.raw), list(self.k)])

    def vec(self) -> np.ndarray:
        """24 doubles in the sfhydro output layout."""
        o = np.zeros(24)
        for k, name in enumerate(("Fb", "Tb", "Fdq", "Tdq", "Fdf", "Tdf")):
            o[3 * k:3 * k + 3] = list(getattr(self, name))
        o[18], o[19], o[20], o[21] = self.Swet, self.Vsub, self.written, self.position
        return o


class orc_body(C.Structure):
    _fields_ = [("mesh", C.POINTER(orc_mesh)), ("T_CG2C", C.c_double * 16), ("T_CG2O", C.c_double * 16),
                ("aabb_min", C.c_double * 3), ("aabb_max", C.c_double * 3), ("Cd", C.c_double * 3),
                ("Cf", C.c_double * 3), ("volume", C.c_double), ("surface", C.c_double), ("P_CB", C.c_double * 3),
                ("buoyant", C.c_int32), ("hydro", C.c_int32)]


class orc_part(C.Structure):
    _fields_ = [("solid", C.POINTER(orc_body)), ("origin", C.c_double * 16), ("O2C", C.c_double * 16),
                ("external", C.c_int32)]


_lib = None


def load(path: str | None = None) -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    p = Path(path) if path else HERE / "liboracle.so"
    if not p.exists():
        raise RuntimeError(f"{p} missing: build the oracle (make -C oracle)")
    lib = C.CDLL(str(p))
    d16 = C.c_double * 16
    sig = {
        "orc_quat_to_gl": (None, [dp, dp, dp]),
        "orc_gl_mul": (None, [dp, dp, dp]),
        "orc_gl_inverse": (None, [dp, dp]),
        "orc_basis_inverse": (None, [dp, dp]),
        "orc_safe_normalize": (None, [dp]),
        "orc_transform_aabb": (None, [dp, dp, C.c_double, dp, dp, dp]),
        "orc_wave_height_64": (C.c_double, [C.POINTER(orc_ocean), C.c_double, C.c_double]),
        "orc_wave_height_32": (C.c_float, [C.POINTER(orc_ocean), C.c_float, C.c_float]),
        "orc_depth_64": (C.c_double, [C.POINTER(orc_ocean), dp]),
        "orc_fluid_velocity": (None, [C.POINTER(orc_ocean), dp, dp]),
        "orc_surface_64": (None, [C.c_int, C.c_int, C.POINTER(orc_mesh), C.POINTER(orc_ocean), dp, dp, dp, dp,
                                  C.POINTER(orc_result)]),
        "orc_surface_32": (None, [C.c_int, C.c_int, C.POINTER(orc_mesh), C.POINTER(orc_ocean), dp, dp, dp, dp,
                                  C.POINTER(orc_result)]),
        "orc_submerged_64": (None, [C.POINTER(orc_mesh), C.POINTER(orc_ocean), dp, dp, dp, dp,
                                    C.POINTER(orc_result)]),
        "orc_submerged_32": (None, [C.POINTER(orc_mesh), C.POINTER(orc_ocean), dp, dp, dp, dp,
                                    C.POINTER(orc_result)]),
        "orc_correct": (None, [C.POINTER(orc_ocean), dp, dp, dp, dp, dp, dp, dp]),
        "orc_classify_64": (C.c_int, [C.POINTER(orc_body), C.POINTER(orc_ocean), dp]),
        "orc_classify_32": (C.c_int, [C.POINTER(orc_body), C.POINTER(orc_ocean), dp]),
        "orc_body_forces_64": (None, [C.POINTER(orc_body), C.POINTER(orc_ocean), dp, dp, dp,
                                      C.POINTER(orc_result)]),
        "orc_body_forces_32": (None, [C.POINTER(orc_body), C.POINTER(orc_ocean), dp, dp, dp,
                                      C.POINTER(orc_result)]),
        "orc_compound_forces_64": (None, [C.POINTER(orc_body), C.POINTER(orc_part), C.c_uint32,
                                          C.POINTER(orc_ocean), dp, dp, dp, C.POINTER(orc_result)]),
        "orc_aero_64": (None, [C.POINTER(orc_mesh), C.c_double, C.POINTER(orc_current), C.c_uint32, dp, dp, dp,
                               dp, dp, dp, dp]),
        "orc_batch_forces_32": (C.c_double, [C.POINTER(orc_body), C.c_uint32, C.POINTER(orc_ocean), dp, C.c_uint32,
                                             C.c_int, C.POINTER(orc_result)]),
        "orc_batch_forces_64": (C.c_double, [C.POINTER(orc_body), C.c_uint32, C.POINTER(orc_ocean), dp, C.c_uint32,
                                             C.c_int, C.POINTER(orc_result)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    del d16
    _lib = lib
    return lib


def _d(a) -> C.Array:
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float64).reshape(-1))
    return (C.c_double * a.size)(*a)


class Mesh:
    """Keeps the numpy buffers alive behind an orc_mesh."""

    def __init__(self, verts: np.ndarray, faces: np.ndarray):
        self.verts = np.ascontiguousarray(verts, dtype=np.float32)
        self.faces = np.ascontiguousarray(faces, dtype=np.uint32)
        self.s = orc_mesh(self.verts.ctypes.data_as(C.POINTER(C.c_float)),
                          self.faces.ctypes.data_as(C.POINTER(C.c_uint32)), self.verts.shape[0],
                          self.faces.shape[0])

    @classmethod
    def from_tri(cls, m) -> "Mesh":
        return cls(m.verts, m.faces)


class Ocean:
    def __init__(self, density=1000.0, gravity=(0.0, 0.0, 9.81), wave_data=None, wave_n=0, channels=1,
                 lengths=(), currents=(), currents_enabled=False):
        self.s = orc_ocean()
        self.s.density = density
        self.s.gravity[:] = list(gravity)
        self.wave = None
        if wave_data is not None:
            self.wave = np.ascontiguousarray(wave_data, dtype=np.float32).reshape(-1)
            self.s.has_waves = 1
            self.s.wave_data = self.wave.ctypes.data_as(C.POINTER(C.c_float))
            self.s.wave_n = wave_n
            self.s.wave_channels = channels
            self.s.wave_layers = len(lengths)
            for i, L in enumerate(lengths):
                self.s.wave_len[i] = L
        self.cur = (orc_current * max(1, len(currents)))()
        for i, (kind, enabled, params) in enumerate(currents):
            self.cur[i].type = kind
            self.cur[i].enabled = int(enabled)
            p = list(params) + [0.0] * (12 - len(params))
            self.cur[i].p[:] = p
        self.s.ncurrents = len(currents)
        self.s.currents = self.cur
        self.s.currents_enabled = int(currents_enabled)


def result(init: np.ndarray | None = None) -> orc_result:
    r = orc_result()
    if init is not None:
        for k, name in enumerate(("Fb", "Tb", "Fdq", "Tdq", "Fdf", "Tdf")):
            getattr(r, name)[:] = list(init[3 * k:3 * k + 3])
        r.Swet, r.Vsub = init[18], init[19]
    return r


def surface(prec: int, damping, buoyancy, mesh: Mesh | None, ocn: Ocean, T_CG, T_C, v, w, init=None,
            full: bool = False):
    """Returns the 24-double output (and with full=True also the 22 extras)."""
    lib = load()
    r = result(init)
    fn = lib.orc_surface_64 if prec == 64 else lib.orc_surface_32
    fn(int(damping), int(buoyancy), C.byref(mesh.s) if mesh is not None else None, C.byref(ocn.s), _d(T_CG),
       _d(T_C), _d(v), _d(w), C.byref(r))
    return (r.vec(), r.extras()) if full else r.vec()


def submerged(prec: int, mesh: Mesh | None, ocn: Ocean, T_CG, T_C, v, w, init=None, full: bool = False):
    lib = load()
    r = result(init)
    fn = lib.orc_submerged_64 if prec == 64 else lib.orc_submerged_32
    fn(C.byref(mesh.s) if mesh is not None else None, C.byref(ocn.s), _d(T_CG), _d(T_C), _d(v), _d(w),
       C.byref(r))
    return (r.vec(), r.extras()) if full else r.vec()


def make_body(mesh: Mesh, T_CG2C=None, T_CG2O=None, aabb_min=None, aabb_max=None, Cd=(1, 1, 1),
              Cf=(0.1, 0.1, 0.1), volume=0.0, surface=0.0, P_CB=(0, 0, 0), buoyant=True, hydro=True) -> orc_body:
    I = np.zeros(16)
    I[[0, 5, 10, 15]] = 1.0
    b = orc_body()
    b.mesh = C.pointer(mesh.s)
    b.T_CG2C[:] = list(I if T_CG2C is None else np.asarray(T_CG2C, dtype=np.float64))
    b.T_CG2O[:] = list(I if T_CG2O is None else np.asarray(T_CG2O, dtype=np.float64))
    if aabb_min is None:
        mn = mesh.verts.min(axis=0).astype(np.float64)
        mx = mesh.verts.max(axis=0).astype(np.float64)
        omin, omax = (C.c_double * 3)(), (C.c_double * 3)()
        load().orc_transform_aabb(_d(mn), _d(mx), 0.0, b.T_CG2C, omin, omax)
        aabb_min, aabb_max = list(omin), list(omax)
    b.aabb_min[:] = list(aabb_min)
    b.aabb_max[:] = list(aabb_max)
    b.Cd[:] = list(Cd)
    b.Cf[:] = list(Cf)
    b.volume = volume
    b.surface = surface
    b.P_CB[:] = list(P_CB)
    b.buoyant = int(buoyant)
    b.hydro = int(hydro)
    b._mesh = mesh   # keep alive
    return b


def body_forces(prec: int, body: orc_body, ocn: Ocean, T_CG, v, w, init=None, full: bool = False):
    lib = load()
    r = result(init)
    fn = lib.orc_body_forces_64 if prec == 64 else lib.orc_body_forces_32
    fn(C.byref(body), C.byref(ocn.s), _d(T_CG), _d(v), _d(w), C.byref(r))
    return (r.vec(), r.extras()) if full else r.vec()


def batch_forces(prec: int, bodies: list, ocn: Ocean, states: np.ndarray, steps: int = 1,
                 nthreads: int = 1) -> tuple[np.ndarray, float]:
    """states: (nbodies, 22); evaluates `steps` passes; returns (out (n,24), seconds)."""
    lib = load()
    n = len(bodies)
    arr = (orc_body * n)(*bodies)
    st = np.ascontiguousarray(states, dtype=np.float64).reshape(n, 22)
    res = (orc_result * n)()
    fn = lib.orc_batch_forces_64 if prec == 64 else lib.orc_batch_forces_32
    secs = fn(arr, n, C.byref(ocn.s), st.ctypes.data_as(dp), steps, nthreads, res)
    return np.stack([r.vec() for r in res]), secs


def correct(ocn: Ocean, Fdq, Tdq, Fdf, Tdf, Cd, Cf, T_O):
    lib = load()
    a, b, c, d = _d(Fdq), _d(Tdq), _d(Fdf), _d(Tdf)
    lib.orc_correct(C.byref(ocn.s), a, b, c, d, _d(Cd), _d(Cf), _d(T_O))
    return np.array(list(a)), np.array(list(b)), np.array(list(c)), np.array(list(d))


def compound_forces(comp: orc_body, parts: list, ocn: Ocean, T_CG, v, w, init=None, full: bool = False):
    """Compound::ComputeHydrodynamicForces (FP64).  parts: list of
    (orc_body solid, origin[16], O2C[16], external)."""
    lib = load()
    arr = (orc_part * max(1, len(parts)))()
    keep = []
    for i, (solid, origin, o2c, external) in enumerate(parts):
        arr[i].solid = C.pointer(solid)
        arr[i].origin[:] = list(np.asarray(origin, dtype=np.float64).reshape(16))
        arr[i].O2C[:] = list(np.asarray(o2c, dtype=np.float64).reshape(16))
        arr[i].external = int(external)
        keep.append(solid)
    r = result(init)
    lib.orc_compound_forces_64(C.byref(comp), arr, len(parts), C.byref(ocn.s), _d(T_CG), _d(v), _d(w), C.byref(r))
    return (r.vec(), r.extras()) if full else r.vec()


def aero(mesh: Mesh | None, gas_density: float, wind, T_CG, T_C, v, w, sums: bool = False):
    """orc_aero_64 (SolidEntity::ComputeAerodynamicForces, SolidEntity.cpp:1865-1918):
    (Fda, Tda); wind: [(kind, params)] summed at every face centroid.  With
    sums, also (S_F, S_T): the sums of the norms of the per-face terms."""
    cur = (orc_current * max(1, len(wind)))()
    for i, (kind, params) in enumerate(wind):
        cur[i].type = kind
        cur[i].enabled = 1
        cur[i].p[:] = list(params) + [0.0] * (12 - len(params))
    F, T, S = (C.c_double * 3)(), (C.c_double * 3)(), (C.c_double * 2)()
    load().orc_aero_64(C.byref(mesh.s) if mesh is not None else None, float(gas_density), cur, len(wind), _d(T_CG),
                       _d(T_C), _d(v), _d(w), F, T, S)
    if sums:
        return np.array(F[:]), np.array(T[:]), (S[0], S[1])
    return np.array(F[:]), np.array(T[:])
