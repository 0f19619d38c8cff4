"""ctypes binding of ``include/sfhydro.h`` (libsfhydro.so).

The library is the only compute path: if it is missing the import fails loudly,
there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIBDIR = Path(__file__).resolve().parent / "_lib"

SFH_OK = 0
SFH_ERR_INVALID = -1
SFH_ERR_SIZE = -2
SFH_ERR_HIP = -3
SFH_ERR_NODEVICE = -4
SFH_ERR_UNSUPPORTED = -5

SFH_STATE_DOUBLES = 22
SFH_OUT_DOUBLES = 24
SFH_NOALLOC_SURFACE = 0
SFH_NOALLOC_SUBMERGED = 1

c_double_p = C.POINTER(C.c_double)
c_float_p = C.POINTER(C.c_float)
c_uint32_p = C.POINTER(C.c_uint32)


class sfh_settings(C.Structure):
    _fields_ = [("damping_forces", C.c_int32), ("realistic_buoyancy", C.c_int32)]


class sfh_body_desc(C.Structure):
    _fields_ = [
        ("mesh", C.c_void_p),
        ("T_CG2C", C.c_double * 16),
        ("T_CG2O", C.c_double * 16),
        ("aabb_min", C.c_double * 3),
        ("aabb_max", C.c_double * 3),
        ("Cd", C.c_double * 3),
        ("Cf", C.c_double * 3),
        ("volume", C.c_double),
        ("surface", C.c_double),
        ("P_CB", C.c_double * 3),
        ("buoyant", C.c_int32),
        ("hydro", C.c_int32),
    ]


class sfh_part_desc(C.Structure):
    _fields_ = [
        ("mesh", C.c_void_p),
        ("origin", C.c_double * 16),
        ("O2C", C.c_double * 16),
        ("Cd", C.c_double * 3),
        ("Cf", C.c_double * 3),
        ("external", C.c_int32),
        ("buoyant", C.c_int32),
        ("hydro", C.c_int32),
    ]


class sfh_compound_desc(C.Structure):
    _fields_ = [("body", sfh_body_desc), ("parts", C.POINTER(sfh_part_desc)), ("nparts", C.c_uint32)]


class sfh_aero_desc(C.Structure):
    _fields_ = [("mesh", C.c_void_p), ("T_CG2C", C.c_double * 16), ("T_CG2H", C.c_double * 16),
                ("approx_params", C.c_double * 3), ("approx_type", C.c_int32), ("aero", C.c_int32)]


class sfh_aero_part_desc(C.Structure):
    _fields_ = [("mesh", C.c_void_p), ("origin", C.c_double * 16), ("O2C", C.c_double * 16),
                ("T_CG2H", C.c_double * 16), ("approx_params", C.c_double * 3), ("approx_type", C.c_int32),
                ("external", C.c_int32)]


class sfh_aero_compound_desc(C.Structure):
    _fields_ = [("T_CG2O", C.c_double * 16), ("aero", C.c_int32), ("pad_", C.c_int32),
                ("parts", C.POINTER(sfh_aero_part_desc)), ("nparts", C.c_uint32)]


class sfh_batch_opts(C.Structure):
    _fields_ = [("faces_per_lane", C.c_uint32), ("waves", C.c_uint32), ("split_passes", C.c_uint32)]


class sfh_tune_point(C.Structure):
    _fields_ = [("faces_per_lane", C.c_uint32), ("waves_per_cu", C.c_uint32), ("split_passes", C.c_uint32),
                ("pad_", C.c_uint32), ("step_ms", C.c_double), ("face_ms", C.c_double)]


class sfh_wave_component(C.Structure):
    _fields_ = [("amplitude", C.c_float), ("phase", C.c_float), ("mx", C.c_int32), ("my", C.c_int32)]


SFH_SUR_FEATURES = 13
SFH_SUR_ROW = 32


class sfh_surrogate_norm(C.Structure):
    _fields_ = [
        ("in_shift", C.c_float * 13),
        ("in_scale", C.c_float * 13),
        ("in_lo", C.c_float * 13),
        ("in_hi", C.c_float * 13),
        ("out_shift", C.c_float * 6),
        ("out_scale", C.c_float * 6),
        ("gate_band", C.c_float),
    ]


class sfh_net_ptrs(C.Structure):
    _fields_ = [("weights", C.c_void_p * 5), ("bias", C.c_void_p * 5)]


class sfh_rigid(C.Structure):
    _fields_ = [("mass", C.c_double), ("inertia", C.c_double * 3)]


class sfh_mesh_properties(C.Structure):
    _fields_ = [("mass", C.c_double), ("volume", C.c_double), ("surface", C.c_double), ("CG", C.c_double * 3),
                ("Ipri", C.c_double * 3), ("Irot", C.c_double * 9)]


SFH_APPROX_AUTO, SFH_APPROX_SPHERE, SFH_APPROX_CYLINDER, SFH_APPROX_ELLIPSOID = 0, 1, 2, 3


class sfh_fluid_approx(C.Structure):
    _fields_ = [("type", C.c_int32), ("iterations", C.c_int32), ("params", C.c_double * 3),
                ("T_CG2H", C.c_double * 16), ("added_mass", C.c_double * 3), ("added_inertia", C.c_double * 3),
                ("Cd", C.c_double * 3), ("Cf", C.c_double * 3)]


# name -> (restype, argtypes)
_PROTOS = {
    "sfh_last_error": (C.c_char_p, []),
    "sfh_init": (C.c_int, [C.c_int]),
    "sfh_mesh_create": (C.c_int, [C.c_void_p, C.c_size_t, C.c_uint32, c_uint32_p, C.c_uint32, C.POINTER(C.c_void_p)]),
    "sfh_mesh_destroy": (C.c_int, [C.c_void_p]),
    "sfh_mesh_info": (C.c_int, [C.c_void_p, c_uint32_p, c_uint32_p, c_float_p, c_float_p]),
    "sfh_ocean_create": (C.c_int, [C.c_double, c_double_p, C.POINTER(C.c_void_p)]),
    "sfh_ocean_destroy": (C.c_int, [C.c_void_p]),
    "sfh_ocean_set_waves": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_uint32, C.c_uint32, C.c_uint32, c_float_p]),
    "sfh_ocean_add_current": (C.c_int, [C.c_void_p, C.c_int, C.c_int, c_double_p]),
    "sfh_ocean_enable_currents": (C.c_int, [C.c_void_p, C.c_int]),
    "sfh_forces_surface": (C.c_int, [C.POINTER(sfh_settings), C.c_void_p, C.c_void_p, c_double_p, c_double_p,
                                     c_double_p, c_double_p, c_double_p, c_double_p, c_double_p, c_double_p,
                                     c_double_p, c_double_p, c_double_p, c_double_p]),
    "sfh_forces_submerged": (C.c_int, [C.c_void_p, C.c_void_p, c_double_p, c_double_p, c_double_p, c_double_p,
                                       c_double_p, c_double_p, c_double_p, c_double_p]),
    "sfh_correct": (C.c_int, [C.c_void_p, c_double_p, c_double_p, c_double_p, c_double_p, c_double_p, c_double_p,
                              c_double_p]),
    "sfh_forces_noalloc": (C.c_int, [C.POINTER(sfh_settings), C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                                     C.c_void_p, C.c_void_p]),
    "sfh_batch_create": (C.c_int, [C.POINTER(sfh_body_desc), C.c_uint32, C.POINTER(sfh_batch_opts),
                                   C.POINTER(C.c_void_p)]),
    "sfh_compound_batch_create": (C.c_int, [C.POINTER(sfh_compound_desc), C.c_uint32, C.POINTER(sfh_batch_opts),
                                            C.POINTER(C.c_void_p)]),
    "sfh_aero_batch_create": (C.c_int, [C.POINTER(sfh_aero_desc), C.c_uint32, C.POINTER(sfh_batch_opts),
                                        C.POINTER(C.c_void_p)]),
    "sfh_aero_compound_batch_create": (C.c_int, [C.POINTER(sfh_aero_compound_desc), C.c_uint32,
                                                 C.POINTER(sfh_batch_opts), C.POINTER(C.c_void_p)]),
    "sfh_tick_create": (C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(C.c_void_p)]),
    "sfh_tick_step": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                C.POINTER(C.c_int), C.c_void_p]),
    "sfh_tick_destroy": (C.c_int, [C.c_void_p]),
    "sfh_wave_spectrum": (C.c_int, [C.c_float, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p]),
    "sfh_wave_fft_create": (C.c_int, [C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(C.c_void_p)]),
    "sfh_wave_fft_step": (C.c_int, [C.c_void_p, C.c_float, C.c_void_p, C.c_void_p]),
    "sfh_wave_fft_info": (C.c_int, [C.c_void_p, c_uint32_p]),
    "sfh_wave_fft_destroy": (C.c_int, [C.c_void_p]),
    "sfh_batch_destroy": (C.c_int, [C.c_void_p]),
    "sfh_batch_set_opts": (C.c_int, [C.c_void_p, C.POINTER(sfh_batch_opts)]),
    "sfh_batch_get_opts": (C.c_int, [C.c_void_p, C.POINTER(sfh_batch_opts)]),
    "sfh_batch_autotune": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                     C.c_uint32, C.POINTER(sfh_tune_point), C.c_uint32, C.POINTER(C.c_uint32)]),
    "sfh_batch_info": (C.c_int, [C.c_void_p, c_uint32_p, c_uint32_p, C.POINTER(C.c_uint64)]),
    "sfh_batch_compute": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "sfh_batch_compute_host": (C.c_int, [C.c_void_p, C.c_void_p, c_double_p, c_double_p]),
    "sfh_batch_compute_masked": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "sfh_batch_capture": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "sfh_batch_replay": (C.c_int, [C.c_void_p, C.c_void_p]),
    "sfh_batch_timing": (C.c_int, [C.c_void_p, C.c_int]),
    "sfh_batch_kernel_ms": (C.c_int, [C.c_void_p, c_double_p, c_double_p, c_uint32_p]),
    "sfh_wave_grid_build": (C.c_int, [C.POINTER(sfh_wave_component), C.c_uint32, C.c_uint32, C.c_void_p,
                                      C.c_void_p]),
    "sfh_surrogate_features": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32,
                                         C.POINTER(sfh_surrogate_norm)] + [C.c_void_p] * 8),
    "sfh_surrogate_set_table": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p, C.c_uint32, C.c_uint32,
                                          C.c_uint32]),
    "sfh_surrogate_plan_create": (C.c_int, [C.c_void_p, C.c_uint32, c_uint32_p, C.POINTER(sfh_surrogate_norm),
                                            C.POINTER(sfh_net_ptrs), C.POINTER(C.c_void_p)]),
    "sfh_surrogate_plan_destroy": (C.c_int, [C.c_void_p]),
    "sfh_surrogate_run": (C.c_int, [C.c_void_p] * 7),
    "sfh_surrogate_merge": (C.c_int, [C.c_void_p] * 6),
    "sfh_forces_aero": (C.c_int, [C.c_void_p, C.c_void_p] + [C.POINTER(C.c_double)] * 6),
    "sfh_correct_aero": (C.c_int, [C.c_int] + [C.POINTER(C.c_double)] * 4),
    "sfh_compute_physical_properties": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32, C.c_double,
                                                  C.c_double, C.POINTER(sfh_mesh_properties)]),
    "sfh_compute_fluid_dynamics_approx": (C.c_int, [C.c_void_p, C.c_uint32, C.c_int, C.POINTER(C.c_double),
                                                    C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_double,
                                                    C.POINTER(sfh_fluid_approx)]),
    "sfh_rollout_integrate": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p,
                                        C.POINTER(sfh_rigid),
                                        C.c_double, C.c_void_p]),
    "sfh_surrogate_apply": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32,
                                      C.POINTER(sfh_surrogate_norm)] + [C.c_void_p] * 7),
}

EXPORTED = tuple(_PROTOS)

_lib = None


def lib_path() -> Path:
    return LIBDIR / "libsfhydro.so"


def load() -> C.CDLL:
    """Loads libsfhydro.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    p = lib_path()
    if not p.exists():
        raise RuntimeError(
            f"{p} is missing: the HIP library has not been built "
            "(run `python -m stonefish_amd.build` or __graft_entry__.build())")
    lib = C.CDLL(str(p))
    for name, (res, args) in _PROTOS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class SfhError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"sfhydro error {code}: {msg}")
        self.code = code


def check(rc: int) -> None:
    if rc != SFH_OK:
        msg = load().sfh_last_error()
        raise SfhError(rc, msg.decode() if msg else "")


# ---- libsfmlp.so (include/sfmlp.h) ------------------------------------------

SFM_OK = 0
SFM_IN = 13
SFM_IN_PAD = 32
SFM_HIDDEN = 256
SFM_LAYERS = 5
SFM_OUT = 6
SFM_OUT_PAD = 32
SFM_DIMS = (SFM_IN_PAD, SFM_HIDDEN, SFM_HIDDEN, SFM_HIDDEN, SFM_HIDDEN, SFM_OUT_PAD)

_MLP_PROTOS = {
    "sfm_last_error": (C.c_char_p, []),
    "sfm_create": (C.c_int, [C.c_uint32, C.c_uint64, C.POINTER(C.c_void_p)]),
    "sfm_destroy": (C.c_int, [C.c_void_p]),
    "sfm_param_count": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint64)]),
    "sfm_layer_io": (C.c_int, [C.c_void_p, C.c_int, c_float_p, c_float_p, C.c_int]),
    "sfm_forward": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p]),
    "sfm_forward_group": (C.c_int, [C.POINTER(C.c_void_p), C.c_uint32, c_uint32_p, C.c_void_p, C.c_void_p,
                                    C.c_void_p]),
    "sfm_forward_layers": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p]),
    "sfm_train_step": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_float, C.c_float, C.c_float,
                                 C.c_float, C.c_void_p, C.c_void_p]),
    "sfm_gradients": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]),
    "sfm_to_bf16": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p]),
    "sfm_probe_mfma": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "sfm_weights": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p)]),
    "sfm_timing": (C.c_int, [C.c_void_p, C.c_int]),
    "sfm_kernel_ms": (C.c_int, [C.c_void_p, c_double_p, c_double_p, c_uint32_p]),
}

MLP_EXPORTED = tuple(_MLP_PROTOS)

_mlp = None


def mlp_lib_path() -> Path:
    return LIBDIR / "libsfmlp.so"


def load_mlp() -> C.CDLL:
    """Loads libsfmlp.so (raises if it has not been built)."""
    global _mlp
    if _mlp is not None:
        return _mlp
    p = mlp_lib_path()
    if not p.exists():
        raise RuntimeError(
            f"{p} is missing: the HIP library has not been built "
            "(run `python -m stonefish_amd.build` or __graft_entry__.build())")
    lib = C.CDLL(str(p))
    for name, (res, args) in _MLP_PROTOS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _mlp = lib
    return lib


class SfmError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"sfmlp error {code}: {msg}")
        self.code = code


def check_mlp(rc: int) -> None:
    if rc != SFM_OK:
        msg = load_mlp().sfm_last_error()
        raise SfmError(rc, msg.decode() if msg else "")
