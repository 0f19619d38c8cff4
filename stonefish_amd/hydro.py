"""Host-side mirror of the reference's hydrodynamics interface, over the C ABI.

Names and argument meaning follow the reference:
  HydrodynamicsSettings             Ocean.h:36-40
  ComputeHydrodynamicForcesSurface  SolidEntity.h:135-137 / SolidEntity.cpp:1289-1721
  ComputeHydrodynamicForcesSubmerged SolidEntity.h:152-153 / SolidEntity.cpp:1723-1797
  CorrectHydrodynamicForces         SolidEntity.h:113-114 / SolidEntity.cpp:1263-1287
  BodyBatch                         the per-body fan-out of SimulationManager.cpp:1619-1655
                                    around SolidEntity::ComputeHydrodynamicForces (1799-1850)

Output vectors are numpy float64 arrays of shape (3,), written in place; the
reference's by-reference scalars (Swet, Vsub) live in a (2,) array.  Outputs the
reference leaves untouched keep their values.  Every computation runs in
libsfhydro.so on the GPU.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _native as N
from .meshes import TriMesh


def _dp(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(N.c_double_p)


def _vec(x, n=3) -> np.ndarray:
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float64).reshape(-1))
    if a.size != n:
        raise ValueError(f"expected {n} values, got {a.size}")
    return a


def init(device: int = 0) -> None:
    N.check(N.load().sfh_init(device))


@dataclass
class HydrodynamicsSettings:
    dampingForces: bool = True
    reallisticBuoyancy: bool = True


class Mesh:
    """Device-resident triangle mesh (the reference's `const Mesh*`)."""

    def __init__(self, mesh: TriMesh):
        lib = N.load()
        v = np.ascontiguousarray(mesh.verts, dtype=np.float32)
        f = np.ascontiguousarray(mesh.faces, dtype=np.uint32)
        h = C.c_void_p()
        N.check(lib.sfh_mesh_create(v.ctypes.data_as(C.c_void_p), 12, v.shape[0],
                                    f.ctypes.data_as(N.c_uint32_p), f.shape[0], C.byref(h)))
        self.handle = h
        self.nv, self.nf = int(v.shape[0]), int(f.shape[0])
        self.aabb_min = v.min(axis=0).astype(np.float64) if v.size else np.zeros(3)
        self.aabb_max = v.max(axis=0).astype(np.float64) if v.size else np.zeros(3)

    def close(self) -> None:
        if self.handle:
            N.check(N.load().sfh_mesh_destroy(self.handle))
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Ocean:
    """Liquid, free surface and currents (the reference's `Ocean*`)."""

    def __init__(self, density: float = 1000.0, gravity=(0.0, 0.0, 9.81)):
        lib = N.load()
        h = C.c_void_p()
        g = _vec(gravity)
        N.check(lib.sfh_ocean_create(float(density), _dp(g), C.byref(h)))
        self.handle = h
        self.density = float(density)
        self.gravity = g

    def set_waves(self, data, n: int, channels: int, lengths, device_ptr: int | None = None) -> None:
        """Wave height texels (n x n x channels, float32); layer l = channel l,
        physical size lengths[l]; None for a flat surface."""
        lib = N.load()
        if data is None and device_ptr is None:
            N.check(lib.sfh_ocean_set_waves(self.handle, None, 1, 0, 0, 0, None))
            return
        L = np.ascontiguousarray(np.asarray(lengths, dtype=np.float32).reshape(-1))
        if device_ptr is not None:
            N.check(lib.sfh_ocean_set_waves(self.handle, C.c_void_p(device_ptr), 0, n, channels, L.size,
                                            L.ctypes.data_as(N.c_float_p)))
        else:
            d = np.ascontiguousarray(data, dtype=np.float32)
            assert d.size == n * n * channels
            N.check(lib.sfh_ocean_set_waves(self.handle, d.ctypes.data_as(C.c_void_p), 1, n, channels, L.size,
                                            L.ctypes.data_as(N.c_float_p)))

    def add_uniform_current(self, v, enabled: bool = True) -> None:
        p = np.zeros(12)
        p[:3] = _vec(v)
        N.check(N.load().sfh_ocean_add_current(self.handle, 0, int(enabled), _dp(p)))

    def add_current(self, kind: int, params, enabled: bool = True) -> None:
        p = np.zeros(12)
        params = np.asarray(params, dtype=np.float64).reshape(-1)
        p[:params.size] = params
        N.check(N.load().sfh_ocean_add_current(self.handle, int(kind), int(enabled), _dp(p)))

    def add_jet(self, point, direction, radius: float, outlet_velocity: float, enabled: bool = True) -> None:
        """Jet(point, direction, radius, outletVelocity) (Jet.cpp:31-37): axis normalised."""
        d = _vec(direction)
        p = np.zeros(12)
        p[:3] = _vec(point)
        p[3:6] = d / np.linalg.norm(d)
        p[6], p[7] = radius, outlet_velocity
        self.add_current(1, p, enabled)

    def add_pipe(self, point1, point2, radius1: float, radius2: float, inlet_velocity: float, exponent: float,
                 enabled: bool = True) -> None:
        """Pipe(point1, point2, radius1, radius2, inletVelocity, exponent) (Pipe.cpp:31-41):
        length |p2 - p1|, axis (p2 - p1) / length."""
        p1, p2 = _vec(point1), _vec(point2)
        ln = float(np.linalg.norm(p2 - p1))
        p = np.zeros(12)
        p[:3] = p1
        p[3:6] = (p2 - p1) / ln
        p[6], p[7], p[8], p[9], p[10] = ln, radius1, radius2, inlet_velocity, exponent
        self.add_current(2, p, enabled)

    def enable_currents(self, on: bool = True) -> None:
        N.check(N.load().sfh_ocean_enable_currents(self.handle, int(on)))

    def close(self) -> None:
        if self.handle:
            N.check(N.load().sfh_ocean_destroy(self.handle))
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Atmosphere(Ocean):
    """Gas and wind fields (the reference's `Atmosphere*`) as seen by the
    aerodynamic face loop: a flat Ocean with the gas density whose currents are
    the wind fields, all of them summed (Atmosphere::GetFluidVelocity,
    Atmosphere.cpp:146-152, checks no enabled flag)."""

    def __init__(self, density: float = 1.225, gravity=(0.0, 0.0, 9.81)):
        super().__init__(density, gravity)
        self.enable_currents(True)

    def add_current(self, kind: int, params, enabled: bool = True) -> None:
        super().add_current(kind, params, True)

    def add_uniform_current(self, v, enabled: bool = True) -> None:
        super().add_uniform_current(v, True)

    def enable_currents(self, on: bool = True) -> None:
        super().enable_currents(True)


def ComputeHydrodynamicForcesSurface(settings: HydrodynamicsSettings, mesh: Mesh | None, ocn: Ocean,
                                     T_CG, T_C, v, omega, Fb, Tb, Fdq, Tdq, Fdf, Tdf, SwetVsub) -> None:
    """SolidEntity::ComputeHydrodynamicForcesSurface (SolidEntity.cpp:1289-1721)."""
    lib = N.load()
    s = N.sfh_settings(int(settings.dampingForces), int(settings.reallisticBuoyancy))
    T_CG, T_C, v, omega = _vec(T_CG, 16), _vec(T_C, 16), _vec(v), _vec(omega)
    sw = np.zeros(1)
    vs = np.zeros(1)
    sw[0], vs[0] = SwetVsub[0], SwetVsub[1]
    N.check(lib.sfh_forces_surface(C.byref(s), mesh.handle if mesh is not None else None, ocn.handle,
                                   _dp(T_CG), _dp(T_C), _dp(v), _dp(omega), _dp(Fb), _dp(Tb), _dp(Fdq), _dp(Tdq),
                                   _dp(Fdf), _dp(Tdf), _dp(sw), _dp(vs)))
    SwetVsub[0], SwetVsub[1] = sw[0], vs[0]


def ComputeHydrodynamicForcesSubmerged(mesh: Mesh | None, ocn: Ocean, T_CG, T_C, v, omega,
                                       Fdq, Tdq, Fdf, Tdf) -> None:
    """SolidEntity::ComputeHydrodynamicForcesSubmerged (SolidEntity.cpp:1723-1797)."""
    lib = N.load()
    T_CG, T_C, v, omega = _vec(T_CG, 16), _vec(T_C, 16), _vec(v), _vec(omega)
    N.check(lib.sfh_forces_submerged(mesh.handle if mesh is not None else None, ocn.handle, _dp(T_CG), _dp(T_C),
                                     _dp(v), _dp(omega), _dp(Fdq), _dp(Tdq), _dp(Fdf), _dp(Tdf)))


def forces_noalloc(settings: HydrodynamicsSettings, mode: int, mesh: Mesh, ocn: Ocean, d_state: int, d_out: int,
                   stream: int = 0) -> None:
    """sfh_forces_noalloc: the per-body Surface (mode N.SFH_NOALLOC_SURFACE) or
    Submerged (N.SFH_NOALLOC_SUBMERGED) evaluation from device memory, with no
    device allocation: d_state points to 38 doubles (T_CG, T_C, v, omega),
    d_out to a 24-double output row; asynchronous on `stream`.  No
    counterpart in the reference (for callers that cannot allocate)."""
    s = N.sfh_settings(int(settings.dampingForces), int(settings.reallisticBuoyancy))
    N.check(N.load().sfh_forces_noalloc(C.byref(s), int(mode), mesh.handle, ocn.handle, C.c_void_p(d_state),
                                        C.c_void_p(d_out), C.c_void_p(stream)))


def ComputeAerodynamicForces(mesh: Mesh | None, atm: Ocean, T_CG, T_C, v, omega, Fda, Tda) -> None:
    """SolidEntity::ComputeAerodynamicForces(mesh, atm, ...) (SolidEntity.cpp:1865-1918).
    atm: an Atmosphere (or an Ocean made with the gas density whose enabled
    currents are the wind fields)."""
    lib = N.load()
    T_CG, T_C, v, omega = _vec(T_CG, 16), _vec(T_C, 16), _vec(v), _vec(omega)
    N.check(lib.sfh_forces_aero(mesh.handle if mesh is not None else None, atm.handle, _dp(T_CG), _dp(T_C), _dp(v),
                                _dp(omega), _dp(Fda), _dp(Tda)))


def CorrectAerodynamicForces(approx_type: int, params, T_CG2H, Fda, Tda) -> None:
    """SolidEntity::CorrectAerodynamicForces (SolidEntity.cpp:1920-1948), in place."""
    p = _vec(params)
    N.check(N.load().sfh_correct_aero(int(approx_type), _dp(p), _dp(_vec(T_CG2H, 16)), _dp(Fda), _dp(Tda)))


def CorrectHydrodynamicForces(ocn: Ocean, Fdq, Tdq, Fdf, Tdf, fdCd, fdCf, T_O) -> None:
    """SolidEntity::CorrectHydrodynamicForces (SolidEntity.cpp:1263-1287)."""
    lib = N.load()
    Cd, Cf, T_O = _vec(fdCd), _vec(fdCf), _vec(T_O, 16)
    N.check(lib.sfh_correct(ocn.handle, _dp(Fdq), _dp(Tdq), _dp(Fdf), _dp(Tdf), _dp(Cd), _dp(Cf), _dp(T_O)))


@dataclass
class BodyDesc:
    """What SolidEntity::ComputeHydrodynamicForces reads from the entity besides
    its state (SolidEntity.h:400-437)."""
    mesh: Mesh
    T_CG2C: np.ndarray = field(default_factory=lambda: _ident())
    T_CG2O: np.ndarray = field(default_factory=lambda: _ident())
    aabb_min: np.ndarray | None = None   # CG-frame collision AABB; default: mesh AABB through T_CG2C
    aabb_max: np.ndarray | None = None
    Cd: np.ndarray = field(default_factory=lambda: np.ones(3))
    Cf: np.ndarray = field(default_factory=lambda: 0.1 * np.ones(3))
    volume: float = 0.0
    surface: float = 0.0
    P_CB: np.ndarray = field(default_factory=lambda: np.zeros(3))
    buoyant: bool = True
    hydro: bool = True


def _ident():
    m = np.zeros(16)
    m[[0, 5, 10, 15]] = 1.0
    return m


def cg_frame_aabb(mesh_min, mesh_max, T_CG2C, margin: float = 0.0):
    """btTransformAabb of the collision child's AABB by T_CG2C (btAabbUtil2.h:182-196):
    the btCompoundShape local AABB of a single-solid body."""
    T = np.asarray(T_CG2C, dtype=np.float64)
    B = np.stack([T[0:3], T[4:7], T[8:11]], axis=1)
    h = 0.5 * (np.asarray(mesh_max) - np.asarray(mesh_min)) + margin
    c = 0.5 * (np.asarray(mesh_max) + np.asarray(mesh_min))
    wc = B @ c + T[12:15]
    e = np.abs(B) @ h
    return wc - e, wc + e


def _fill_body(d, b: BodyDesc) -> None:
    d.mesh = b.mesh.handle if b.mesh is not None else None
    d.T_CG2C[:] = list(_vec(b.T_CG2C, 16))
    d.T_CG2O[:] = list(_vec(b.T_CG2O, 16))
    if b.aabb_min is None:
        mn, mx = cg_frame_aabb(b.mesh.aabb_min, b.mesh.aabb_max, b.T_CG2C)
    else:
        mn, mx = b.aabb_min, b.aabb_max
    d.aabb_min[:] = list(_vec(mn))
    d.aabb_max[:] = list(_vec(mx))
    d.Cd[:] = list(_vec(b.Cd))
    d.Cf[:] = list(_vec(b.Cf))
    d.volume = float(b.volume)
    d.surface = float(b.surface)
    d.P_CB[:] = list(_vec(b.P_CB))
    d.buoyant = int(b.buoyant)
    d.hydro = int(b.hydro)


@dataclass
class PartDesc:
    """One part of a compound (Compound.h CompoundPart and the part solid)."""
    mesh: Mesh
    origin: np.ndarray = field(default_factory=lambda: _ident())
    O2C: np.ndarray = field(default_factory=lambda: _ident())
    Cd: np.ndarray = field(default_factory=lambda: np.ones(3))
    Cf: np.ndarray = field(default_factory=lambda: 0.1 * np.ones(3))
    external: bool = True
    buoyant: bool = True
    hydro: bool = True


@dataclass
class CompoundDesc:
    """A compound body (Compound::ComputeHydrodynamicForces, Compound.cpp:291-437):
    `body` carries T_CG2O, the CG-frame AABB (required), volume, surface, P_CB,
    buoyant, hydro; its mesh and T_CG2C are unused."""
    body: BodyDesc
    parts: list[PartDesc]


class BodyBatch:
    """Many bodies evaluated in one pass (SolidEntity::ComputeHydrodynamicForces
    for each, with Ocean::ApplyFluidForces' settings {damping, buoyancy} = {true, true})."""

    def __init__(self, bodies: list[BodyDesc], faces_per_lane: int = 0, waves: int = 0, split_passes: int = 0):
        lib = N.load()
        arr = (N.sfh_body_desc * max(1, len(bodies)))()
        for i, b in enumerate(bodies):
            _fill_body(arr[i], b)
        opts = N.sfh_batch_opts(faces_per_lane, waves, split_passes)
        h = C.c_void_p()
        N.check(lib.sfh_batch_create(arr, len(bodies), C.byref(opts), C.byref(h)))
        self.handle = h
        self.n = len(bodies)
        self.bodies = bodies   # keep meshes alive

    def compute_masked(self, ocn: Ocean, d_states: int, d_active: int | None, d_out: int, stream: int = 0) -> None:
        """sfh_batch_compute_masked: only rows with d_active[i] != 0 (device bytes) are evaluated and written."""
        N.check(N.load().sfh_batch_compute_masked(self.handle, ocn.handle, C.c_void_p(d_states),
                                                  C.c_void_p(d_active or 0), C.c_void_p(d_out),
                                                  C.c_void_p(stream)))

    def info(self):
        nb, nc, nf = C.c_uint32(), C.c_uint32(), C.c_uint64()
        N.check(N.load().sfh_batch_info(self.handle, C.byref(nb), C.byref(nc), C.byref(nf)))
        return nb.value, nc.value, nf.value

    def set_opts(self, faces_per_lane: int = 0, waves: int = 0, split_passes: int = 0) -> None:
        opts = N.sfh_batch_opts(faces_per_lane, waves, split_passes)
        N.check(N.load().sfh_batch_set_opts(self.handle, C.byref(opts)))

    def get_opts(self) -> tuple[int, int, int]:
        """(faces_per_lane, waves, split_passes) the batch runs with."""
        opts = N.sfh_batch_opts()
        N.check(N.load().sfh_batch_get_opts(self.handle, C.byref(opts)))
        return int(opts.faces_per_lane), int(opts.waves), int(opts.split_passes)

    def autotune(self, ocn: Ocean, d_states: int, d_out: int, stream: int = 0, d_active: int | None = None,
                 reps: int = 5) -> list[dict]:
        """sfh_batch_autotune: times the step over faces_per_lane x waves per
        CU x split_passes on these inputs and keeps the fastest; returns the
        measured points."""
        pts = (N.sfh_tune_point * 128)()
        n = C.c_uint32()
        N.check(N.load().sfh_batch_autotune(self.handle, ocn.handle, C.c_void_p(d_states), C.c_void_p(d_active or 0),
                                            C.c_void_p(d_out), C.c_void_p(stream), reps, pts, 128, C.byref(n)))
        return [{"faces_per_lane": p.faces_per_lane, "waves_per_cu": p.waves_per_cu, "split_passes": p.split_passes,
                 "step_ms": round(p.step_ms, 5), "face_ms": round(p.face_ms, 5)} for p in pts[:n.value]]

    def compute_host(self, ocn: Ocean, states: np.ndarray, out: np.ndarray | None = None) -> np.ndarray:
        """states: (n, 22) float64 = T_CG[16], v[3], omega[3]; returns (n, 24)."""
        st = np.ascontiguousarray(states, dtype=np.float64).reshape(self.n, N.SFH_STATE_DOUBLES)
        if out is None:
            out = np.zeros((self.n, N.SFH_OUT_DOUBLES))
        N.check(N.load().sfh_batch_compute_host(self.handle, ocn.handle, _dp(st), _dp(out)))
        return out

    def compute(self, ocn: Ocean, d_states: int, d_out: int, stream: int = 0) -> None:
        """Device pointers (e.g. torch tensor data_ptr()); asynchronous on `stream`."""
        N.check(N.load().sfh_batch_compute(self.handle, ocn.handle, C.c_void_p(d_states), C.c_void_p(d_out),
                                           C.c_void_p(stream)))

    def capture(self, ocn: Ocean, d_states: int, d_out: int, stream: int) -> None:
        N.check(N.load().sfh_batch_capture(self.handle, ocn.handle, C.c_void_p(d_states), C.c_void_p(d_out),
                                           C.c_void_p(stream)))

    def replay(self, stream: int) -> None:
        N.check(N.load().sfh_batch_replay(self.handle, C.c_void_p(stream)))

    def timing(self, on: bool) -> None:
        N.check(N.load().sfh_batch_timing(self.handle, int(on)))

    def kernel_ms(self):
        f, t, c = C.c_double(), C.c_double(), C.c_uint32()
        N.check(N.load().sfh_batch_kernel_ms(self.handle, C.byref(f), C.byref(t), C.byref(c)))
        return f.value, t.value, c.value

    def close(self) -> None:
        if self.handle:
            N.check(N.load().sfh_batch_destroy(self.handle))
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


OUT_FIELDS = {"Fb": slice(0, 3), "Tb": slice(3, 6), "Fdq": slice(6, 9), "Tdq": slice(9, 12),
              "Fdf": slice(12, 15), "Tdf": slice(15, 18), "Swet": 18, "Vsub": 19, "written": 20, "position": 21}


class CompoundBatch(BodyBatch):
    """A batch of compounds: states and outputs per compound; the parts run
    through the face passes as bodies (sfh_compound_batch_create)."""

    def __init__(self, comps: list[CompoundDesc], faces_per_lane: int = 0, waves: int = 0):
        lib = N.load()
        arr = (N.sfh_compound_desc * max(1, len(comps)))()
        self._parts = []
        for i, c in enumerate(comps):
            if c.body.aabb_min is None or c.body.aabb_max is None:
                raise ValueError("a compound needs its CG-frame AABB")
            _fill_body(arr[i].body, c.body)
            parts = (N.sfh_part_desc * max(1, len(c.parts)))()
            for j, p in enumerate(c.parts):
                q = parts[j]
                q.mesh = p.mesh.handle
                q.origin[:] = list(_vec(p.origin, 16))
                q.O2C[:] = list(_vec(p.O2C, 16))
                q.Cd[:] = list(_vec(p.Cd))
                q.Cf[:] = list(_vec(p.Cf))
                q.external = int(p.external)
                q.buoyant = int(p.buoyant)
                q.hydro = int(p.hydro)
            self._parts.append(parts)
            arr[i].parts = parts
            arr[i].nparts = len(c.parts)
        opts = N.sfh_batch_opts(faces_per_lane, waves)
        h = C.c_void_p()
        N.check(lib.sfh_compound_batch_create(arr, len(comps), C.byref(opts), C.byref(h)))
        self.handle = h
        self.n = len(comps)
        self.bodies = comps


@dataclass
class AeroDesc:
    """What SolidEntity::ComputeAerodynamicForces(atm) and
    CorrectAerodynamicForces read from the entity (SolidEntity.cpp:1852-1948)."""
    mesh: Mesh | None
    T_CG2C: np.ndarray = field(default_factory=lambda: _ident())
    T_CG2H: np.ndarray = field(default_factory=lambda: _ident())
    approx_params: np.ndarray = field(default_factory=lambda: np.ones(3))
    approx_type: int = N.SFH_APPROX_SPHERE
    aero: bool = True


@dataclass
class AeroPartDesc:
    """One part of Compound::ComputeAerodynamicForces (Compound.cpp:439-464)."""
    mesh: Mesh
    origin: np.ndarray = field(default_factory=lambda: _ident())
    O2C: np.ndarray = field(default_factory=lambda: _ident())
    T_CG2H: np.ndarray = field(default_factory=lambda: _ident())
    approx_params: np.ndarray = field(default_factory=lambda: np.ones(3))
    approx_type: int = N.SFH_APPROX_SPHERE
    external: bool = True


@dataclass
class AeroCompoundDesc:
    T_CG2O: np.ndarray
    parts: list[AeroPartDesc]
    aero: bool = True


class AeroBatch(BodyBatch):
    """Bodies in the air (sfh_aero_batch_create): compute on an Atmosphere;
    each row holds Fda in [6:9] and Tda in [9:12] after CorrectAerodynamicForces."""

    def __init__(self, bodies: list[AeroDesc], faces_per_lane: int = 0, waves: int = 0):
        lib = N.load()
        arr = (N.sfh_aero_desc * max(1, len(bodies)))()
        for i, b in enumerate(bodies):
            d = arr[i]
            d.mesh = b.mesh.handle if b.mesh is not None else None
            d.T_CG2C[:] = list(_vec(b.T_CG2C, 16))
            d.T_CG2H[:] = list(_vec(b.T_CG2H, 16))
            d.approx_params[:] = list(_vec(b.approx_params))
            d.approx_type = int(b.approx_type)
            d.aero = int(b.aero)
        opts = N.sfh_batch_opts(faces_per_lane, waves)
        h = C.c_void_p()
        N.check(lib.sfh_aero_batch_create(arr, len(bodies), C.byref(opts), C.byref(h)))
        self.handle = h
        self.n = len(bodies)
        self.bodies = bodies


class AeroCompoundBatch(BodyBatch):
    """Compounds in the air (sfh_aero_compound_batch_create); rows per compound."""

    def __init__(self, comps: list[AeroCompoundDesc], faces_per_lane: int = 0, waves: int = 0):
        lib = N.load()
        arr = (N.sfh_aero_compound_desc * max(1, len(comps)))()
        self._parts = []
        for i, c in enumerate(comps):
            arr[i].T_CG2O[:] = list(_vec(c.T_CG2O, 16))
            arr[i].aero = int(c.aero)
            parts = (N.sfh_aero_part_desc * max(1, len(c.parts)))()
            for j, p in enumerate(c.parts):
                q = parts[j]
                q.mesh = p.mesh.handle
                q.origin[:] = list(_vec(p.origin, 16))
                q.O2C[:] = list(_vec(p.O2C, 16))
                q.T_CG2H[:] = list(_vec(p.T_CG2H, 16))
                q.approx_params[:] = list(_vec(p.approx_params))
                q.approx_type = int(p.approx_type)
                q.external = int(p.external)
            self._parts.append(parts)
            arr[i].parts = parts
            arr[i].nparts = len(c.parts)
        opts = N.sfh_batch_opts(faces_per_lane, waves)
        h = C.c_void_p()
        N.check(lib.sfh_aero_compound_batch_create(arr, len(comps), C.byref(opts), C.byref(h)))
        self.handle = h
        self.n = len(comps)
        self.bodies = comps


class TickDriver:
    """sfh_tick_*: the per-tick force step of the simulation loop for every
    body of a batch, recomputing every `prescaler` ticks (fdPrescaler) and
    applying the held forces every tick (SimulationManager.cpp:1583-1655)."""

    def __init__(self, batch: BodyBatch, prescaler: int = 1):
        h = C.c_void_p()
        N.check(N.load().sfh_tick_create(batch.handle, int(prescaler), C.byref(h)))
        self.handle = h
        self.batch = batch

    def step(self, fluid: Ocean, d_states: int, d_overlap: int | None, d_out: int, d_applied: int,
             stream: int = 0) -> bool:
        """Device pointers; returns whether this tick recomputed."""
        r = C.c_int()
        N.check(N.load().sfh_tick_step(self.handle, fluid.handle, C.c_void_p(d_states), C.c_void_p(d_overlap or 0),
                                       C.c_void_p(d_out), C.c_void_p(d_applied), C.byref(r), C.c_void_p(stream)))
        return bool(r.value)

    def close(self) -> None:
        if self.handle:
            N.check(N.load().sfh_tick_destroy(self.handle))
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---- hull preprocessing (host) -------------------------------------------------

def compute_physical_properties(verts, faces, thickness: float = 0.0, density: float = 1000.0) -> dict:
    """ComputePhysicalProperties (GeometryFileUtil.cpp:335-563) through
    sfh_compute_physical_properties: mass, volume, surface, CG, principal
    moments Ipri and axes Irot (columns) of a closed mesh (thickness <= 0) or
    a shell."""
    v = np.ascontiguousarray(verts, dtype=np.float32).reshape(-1, 3)
    f = np.ascontiguousarray(faces, dtype=np.uint32).reshape(-1, 3)
    out = N.sfh_mesh_properties()
    N.check(N.load().sfh_compute_physical_properties(v.ctypes.data, v.shape[0], f.ctypes.data, f.shape[0],
                                                     float(thickness), float(density), C.byref(out)))
    return {"mass": out.mass, "volume": out.volume, "surface": out.surface, "CG": np.array(out.CG[:]),
            "Ipri": np.array(out.Ipri[:]), "Irot": np.array(out.Irot[:]).reshape(3, 3)}


def compute_fluid_dynamics_approx(verts, kind: int = N.SFH_APPROX_AUTO, T_CG2C=None, P_CB=(0.0, 0.0, 0.0),
                                  T_CG2O=None, rho: float = 1000.0) -> dict:
    """SolidEntity::ComputeFluidDynamicsApprox (SolidEntity.cpp:698-1039)
    through sfh_compute_fluid_dynamics_approx: the proxy (sphere, cylinder or
    ellipsoid), its frame, added mass and inertia, and Cd, Cf (origin frame)."""
    v = np.ascontiguousarray(verts, dtype=np.float32).reshape(-1, 3)
    eye = np.eye(4).reshape(-1)
    a = np.ascontiguousarray(T_CG2C if T_CG2C is not None else eye, dtype=np.float64)
    o = np.ascontiguousarray(T_CG2O if T_CG2O is not None else eye, dtype=np.float64)
    out = N.sfh_fluid_approx()
    N.check(N.load().sfh_compute_fluid_dynamics_approx(v.ctypes.data, v.shape[0], int(kind), _dp(a),
                                                       _dp(_vec(P_CB)), _dp(o), float(rho), C.byref(out)))
    return {"type": out.type, "iterations": out.iterations, "params": np.array(out.params[:]),
            "T_CG2H": np.array(out.T_CG2H[:]), "added_mass": np.array(out.added_mass[:]),
            "added_inertia": np.array(out.added_inertia[:]), "Cd": np.array(out.Cd[:]), "Cf": np.array(out.Cf[:])}
