"""ctypes declarations of libsph_hip (include/sph.h, include/sph_batch.h,
include/sph_setup.h).

The library is built in-tree (``make -C smoothedparticles.jl_amd/csrc``) and
loaded from this directory.  There is no fallback: if the library is missing
or fails to load, every entry point raises.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsph_hip.so")

# error codes
SPH_OK = 0
SPH_EINVAL = -1
SPH_EHIP = -2
SPH_ESTATE = -3
SPH_ENOMEM = -4
SPH_EUNSUP = -5

SPH_FP64 = 0
SPH_MIXED = 1

FIELDS = {"x": 0, "v": 1, "a": 2, "rho": 3, "P": 4, "Drho": 5, "rho0": 6, "type": 7}
VECTOR_FIELDS = {"x", "v", "a"}
SPH_FIELD_KEY = 8  # read only: find_key at the last build (1-based)
SPH_SNAPSHOT_MAX_FIELDS = 9

# operator ids
SPH_OP_DENSITY_SUM = 1
SPH_OP_DENSITY_SUM_RHO0 = 2
SPH_OP_INTERNAL_FORCE = 3
SPH_OP_PRESSURE_FORCE_RHO0 = 4
SPH_OP_BALANCE_OF_MASS = 5
SPH_OP_RESET_RHO = 16
SPH_OP_RESET_A = 17
SPH_OP_FIND_PRESSURE = 18
SPH_OP_FIND_PRESSURE_RHO0 = 19
SPH_OP_FIND_PRESSURE_CONTINUITY = 20
SPH_OP_ACCELERATE = 21
SPH_OP_MOVE = 22
SPH_OP_DRIFT = 23
SPH_OP_KICK = 24

SPH_REDUCE_ENERGY = 0
SPH_REDUCE_ENERGY_COLLISION = 1
SPH_REDUCE_MOMENTUM = 2
SPH_REDUCE_ENERGY_RHO0 = 3

SPH_SCHEME_WCSPH_SE = 0
SPH_SCHEME_WCSPH_SE_RHO0 = 1
SPH_SCHEME_COLLAPSE_DRY = 2

SPH_OPTION_PAIR_MASKS = 0
SPH_OPTION_KEPT_LISTS = 1
SPH_OPTION_GRAPH = 2
SPH_PROBE_FLUID_ONLY = 1
SPH_PROBE_NOALLOC_MAX = 65536
SPH_SUM_LJ_WALL = 1

SPH_KERNEL_BUILD = 0
SPH_KERNEL_DENSITY = 1
SPH_KERNEL_FORCE = 2
SPH_KERNEL_UNARY = 3


class sph_params(ctypes.Structure):
    _fields_ = [
        ("m", ctypes.c_double),
        ("rho0", ctypes.c_double),
        ("c", ctypes.c_double),
        ("alpha", ctypes.c_double),
        ("hs", ctypes.c_double),
        ("mu", ctypes.c_double),
        ("nu", ctypes.c_double),
        ("g", ctypes.c_double * 3),
        ("dt", ctypes.c_double),
        ("coeff", ctypes.c_double),
    ]


class sph_system_desc(ctypes.Structure):
    _fields_ = [
        ("h", ctypes.c_double),
        ("bbox_min", ctypes.c_double * 3),
        ("bbox_max", ctypes.c_double * 3),
        ("periodic", ctypes.c_int),
        ("pad", ctypes.c_int),
        ("params", sph_params),
    ]


class sph_lattice_spec(ctypes.Structure):
    _fields_ = [
        ("dim", ctypes.c_int),
        ("kind", ctypes.c_int),
        ("dr", ctypes.c_double),
        ("lo", ctypes.c_double * 3),
        ("hi", ctypes.c_double * 3),
        ("width", ctypes.c_double),
        ("lid_axis", ctypes.c_int),
        ("hydro_axis", ctypes.c_int),
        ("lid", ctypes.c_double),
        ("type", ctypes.c_double),
        ("rho", ctypes.c_double),
        ("P", ctypes.c_double),
        ("rho0", ctypes.c_double),
        ("g", ctypes.c_double),
        ("H", ctypes.c_double),
        ("c", ctypes.c_double),
        ("jitter", ctypes.c_double),
        ("seed", ctypes.c_uint64),
    ]


SPH_LATTICE_BOX = 0
SPH_LATTICE_WALL_LAYER = 1

_d = ctypes.c_double
_i = ctypes.c_int
_i64 = ctypes.c_int64
_p = ctypes.c_void_p
_pd = ctypes.POINTER(ctypes.c_double)
_pi64 = ctypes.POINTER(ctypes.c_int64)

# name -> (restype, argtypes); every function declared in include/*.h
SIGNATURES = {
    "sph_create": (_i, [ctypes.POINTER(_p), _d, _pd, _pd, _i]),
    "sph_destroy": (_i, [_p]),
    "sph_append": (_i, [_p, _i64, _pd, _pd, _pd, _pd, _pd]),
    "sph_count": (_i, [_p, _pi64]),
    "sph_set_field": (_i, [_p, _i, _pd, _i64]),
    "sph_get_field": (_i, [_p, _i, _pd, _i64]),
    "sph_build_cells": (_i, [_p, _pi64]),
    "sph_apply_binary": (_i, [_p, _i, ctypes.POINTER(sph_params), _i]),
    "sph_apply_unary": (_i, [_p, _i, ctypes.POINTER(sph_params)]),
    "sph_apply": (_i, [_p, _i, ctypes.POINTER(sph_params), _i]),
    "sph_reduce": (_i, [_p, _i, ctypes.POINTER(sph_params), _pd]),
    "sph_neighbour_pairs": (_i, [_p, _pi64, _i64, _pi64]),
    "sph_neighbour_digest": (_i, [_p, _pi64, ctypes.POINTER(ctypes.c_uint64), _i64]),
    "sph_probe_sum": (_i, [_p, ctypes.POINTER(sph_params), _i, _pd, _i64, _pd]),
    "sph_batch_probe_sum": (_i, [_p, _i, _i, _pd, _i64, _pd]),
    "sph_probe_sum_noalloc": (_i, [_p, ctypes.POINTER(sph_params), _i, _pd, _i64, _pd]),
    "sph_particle_sum": (_i, [_p, _i, _pd, _i, _pd, _i64]),
    "sph_step": (_i, [_p, _i, ctypes.POINTER(sph_params), _i, _pi64]),
    "sph_step_pairs": (_i, [_p, _pi64, _pi64]),
    "sph_step_tests": (_i, [_p, _pi64]),
    "sph_key_geometry": (_i, [_p, _pi64, _pi64, _pi64, _pi64, ctypes.POINTER(_i)]),
    "sph_set_tile": (_i, [_p, _i]),
    "sph_autotune": (_i, [_p, ctypes.POINTER(sph_params), ctypes.POINTER(_i)]),
    "sph_candidates": (_i, [_p, _pi64, _pi64]),
    "sph_set_option": (_i, [_p, _i, _i]),
    "sph_profile": (_i, [_p, _i]),
    "sph_profile_read": (_i, [_p, _i, _pd, _pi64]),
    "sph_last_error": (ctypes.c_char_p, []),
    "sph_batch_create": (_i, [ctypes.POINTER(_p), _i]),
    "sph_batch_destroy": (_i, [_p]),
    "sph_batch_add": (_i, [_p, ctypes.POINTER(sph_system_desc), _i64, _pd, _pd, _pd, _pd, _pd,
                           ctypes.POINTER(_i)]),
    "sph_batch_retire": (_i, [_p, _i]),
    "sph_batch_size": (_i, [_p, ctypes.POINTER(_i), _pi64]),
    "sph_batch_count": (_i, [_p, _i, _pi64]),
    "sph_batch_get_field": (_i, [_p, _i, _i, _pd, _i64]),
    "sph_batch_set_field": (_i, [_p, _i, _i, _pd, _i64]),
    "sph_batch_reduce": (_i, [_p, _i, _i, _pd]),
    "sph_batch_build_cells": (_i, [_p]),
    "sph_batch_neighbour_pairs": (_i, [_p, _i, _pi64, _i64, _pi64]),
    "sph_batch_neighbour_digest": (_i, [_p, _i, _pi64, ctypes.POINTER(ctypes.c_uint64), _i64]),
    "sph_batch_init_rho0": (_i, [_p]),
    "sph_batch_step": (_i, [_p, _i, _i, _pi64]),
    "sph_batch_set_tile": (_i, [_p, _i]),
    "sph_batch_autotune": (_i, [_p, ctypes.POINTER(_i)]),
    "sph_batch_footprint": (_i, [_p, _pi64]),
    "sph_lattice_box": (_i64, [_i, _d, _pd, _pd, _pd, _i64]),
    "sph_lattice_wall_layer": (_i64, [_i, _d, _pd, _pd, _d, _i, _d, _pd, _i64]),
    "sph_jitter": (None, [_pd, _i64, _i, _d, ctypes.c_uint64]),
    "sph_generate_lattice": (_i, [_p, ctypes.POINTER(sph_lattice_spec), _pi64]),
    "sph_snapshot_begin": (_i, [_p, ctypes.POINTER(_i), _i, ctypes.POINTER(_p)]),
    "sph_batch_snapshot_begin": (_i, [_p, _i, ctypes.POINTER(_i), _i, ctypes.POINTER(_p)]),
    "sph_snapshot_wait": (_i, [_p, _pi64]),
    "sph_snapshot_field": (_i, [_p, _i, ctypes.POINTER(_pd)]),
    "sph_snapshot_free": (_i, [_p]),
    "sph_batch_generate_lattice": (_i, [_p, _i, ctypes.POINTER(sph_lattice_spec), _pi64]),
}

_LIB = None


class SPHError(RuntimeError):
    """A libsph_hip call returned an error code."""

    def __init__(self, code, message):
        super().__init__(f"libsph_hip error {code}: {message}")
        self.code = code


def lib():
    """Load libsph_hip.so (in-tree build).  Raises if it is not built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is not built; run __graft_entry__.build() or "
                "make -C smoothedparticles.jl_amd/csrc"
            )
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def check(rc):
    if rc != 0:
        msg = lib().sph_last_error()
        raise SPHError(rc, msg.decode() if msg else "")
    return rc
