"""ctypes binding of libmpm_hip.so (the C ABI declared in include/mpm_hip.h).

The library is built in-tree by mpm-engine_amd/build.py. Loading fails loudly when it is
missing: there is no fallback path, the HIP kernels are the only implementation.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_PKG_ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
LIB_PATH = os.path.join(_PKG_ROOT, "lib", "libmpm_hip.so")

MAX_MATERIALS = 16

# mode / flag bits, mirrored from include/mpm_hip.h
PASS_G2P, PASS_STRESS, PASS_P2G, PASS_LEAN, PASS_KEEP_STRESS, PASS_VISCOUS_ONLY = 1, 2, 4, 8, 16, 32
PASS_WG128, PASS_PPT2, PASS_PPT4 = 64, 128, 256  # launch shape (results do not depend on it)
PASS_DETERMINISTIC = 512
PASS_PAIR = 1024  # two particles per lane over chunks of up to 512 (deterministic lean pass)
VISCOUS_IDS = frozenset((6, 7, 8, 10, 12))
GRID_ZERO_IN, GRID_FIXED_IN, GRID_WRENCH_DEFER = 1, 2, 4
MAX_FIXED_ORDER_SLOTS = 64  # wrench slots reduced in fixed order; more use float atomics
ERR_STENCIL, ERR_NONFINITE, ERR_SOLID_IN_VISCOUS, ERR_SCENE = 1, 2, 4, 8
ERR_ADJ_MATERIAL, ERR_ADJ_COLLIDER, ERR_FIXED_RANGE, ERR_POOL = 16, 32, 64, 512
COLLIDER_PLANE, COLLIDER_BOX, COLLIDER_CUP, COLLIDER_WALLS, COLLIDER_SDF, COLLIDER_POINTCLOUD = \
    1, 2, 3, 4, 5, 6

_FP = C.c_void_p


class Particles(C.Structure):
    _fields_ = [("n", C.c_int32), ("stride", C.c_int32),
                ("x", _FP), ("v", _FP), ("F", _FP), ("F_trial", _FP), ("C", _FP),
                ("stress", _FP), ("L", _FP), ("vol", _FP), ("mass", _FP), ("mu", _FP),
                ("lam", _FP), ("yield_stress", _FP), ("Jp", _FP), ("material", _FP),
                ("selection", _FP), ("cov", _FP), ("scene", _FP), ("cdf_tag", _FP),
                ("rigid_id", _FP), ("x_ref", _FP), ("modifier_mask", _FP),
                ("chunk_start", _FP), ("chunk_count", _FP), ("chunk_capacity", C.c_int32), ("pad_", C.c_int32)]


# mpm_rigid_body_t / mpm_rigid_plane_t as numpy record types (device arrays of them)
RIGID_BODY_DTYPE = np.dtype([("x_cm", "<f4", 3), ("v_cm", "<f4", 3), ("omega", "<f4", 3),
                             ("R", "<f4", 9), ("mass", "<f4"), ("inv_inertia", "<f4", 9),
                             ("pad_", "<f4", 4)])
RIGID_PLANE_DTYPE = np.dtype([("point", "<f4", 3), ("normal", "<f4", 3), ("restitution", "<f4"),
                              ("friction", "<f4"), ("start_time", "<f4"), ("end_time", "<f4")])
# mpm_particle_modifier_t
MODIFIER_DTYPE = np.dtype([("kind", "<i4"), ("bit", "<i4"), ("start_time", "<f4"), ("end_time", "<f4"),
                           ("point", "<f4", 3), ("vec", "<f4", 3), ("axis1", "<f4", 3), ("axis2", "<f4", 3),
                           ("size", "<f4", 3), ("half_height", "<f4"), ("radius", "<f4"),
                           ("rotation_scale", "<f4"), ("translation_scale", "<f4"), ("pad_", "<f4", 3)])
MAX_MODIFIERS = 32


class Adjoint(C.Structure):
    _fields_ = [(nm, _FP) for nm in ("x", "v", "C", "F_trial", "F", "stress", "mu", "lam",
                                     "yield_stress", "grid_v", "grid_mv", "L", "visc", "bulk",
                                     "hb_K", "hb_n", "mui_mu_s", "mui_dmu", "mui_I0", "mui_d")]


MAX_CDF = 4
CDF_OWNER_SHIFT = 8


class Cdf(C.Structure):
    _fields_ = [("n_cdf", C.c_int32), ("mask_off", C.c_int32), ("tag", _FP), ("tag_prev", _FP),
                ("dist", _FP), ("dist_prev", _FP), ("block", _FP), ("lanes", _FP),
                ("reaction", _FP)]


class Grid(C.Structure):
    _fields_ = [("node_in", _FP), ("node_out", _FP), ("n", C.c_int32 * 3),
                ("nb", C.c_int32 * 3), ("dx", C.c_float), ("inv_dx", C.c_float),
                ("error_flag", _FP), ("n_scenes", C.c_int32), ("pad_", C.c_int32),
                ("cdf", Cdf), ("node_fix", _FP), ("fix_scale", C.c_float * 2), ("block_touch", _FP),
                ("block_slot", _FP), ("pool_blocks", C.c_int32), ("pad2_", C.c_int32)]


_L3 = (np.float32, (MAX_CDF, 3))
CDF_LANES_DTYPE = np.dtype([
    ("center",) + _L3, ("velocity",) + _L3, ("omega",) + _L3,
    ("center_prev",) + _L3, ("velocity_prev",) + _L3, ("omega_prev",) + _L3,
    ("friction", np.float32, (MAX_CDF,)), ("type", np.int32, (MAX_CDF,))])


class CdfCollider(C.Structure):
    _fields_ = [("val", _FP), ("valid", _FP), ("res", C.c_int32), ("lane", C.c_int32),
                ("origin", C.c_float * 3), ("cell", C.c_float), ("center", C.c_float * 3),
                ("quat", C.c_float * 4), ("band", C.c_float), ("time", C.c_float),
                ("start_time", C.c_float), ("end_time", C.c_float)]


_M16 = (np.float32, (MAX_MATERIALS,))
MODEL_DTYPE = np.dtype([
    ("gravity", np.float32, (3,)), ("rpic_damping", np.float32),
    ("grid_v_damping_scale", np.float32), ("update_cov_with_F", np.int32),
    ("periodic_x", np.int32), ("eta_table_n", np.int32), ("eta_table_smin", np.float32),
    ("eta_table_smax", np.float32), ("eta_table", np.uint64),
    ("E",) + _M16, ("nu",) + _M16, ("bulk",) + _M16, ("hardening",) + _M16, ("xi",) + _M16,
    ("plastic_viscosity",) + _M16, ("softening",) + _M16, ("alpha",) + _M16,
    ("muI_mu_s",) + _M16, ("muI_delta_mu",) + _M16, ("muI_I0",) + _M16, ("muI_d",) + _M16,
    ("muI_rho_s",) + _M16, ("muI_phi_init",) + _M16, ("muI_phi_chi",) + _M16,
], align=True)

COLLIDER_DTYPE = np.dtype([
    ("kind", np.int32), ("surface", np.int32), ("reset", np.int32), ("slot", np.int32),
    ("lo", np.int32, (3,)), ("hi", np.int32, (3,)), ("active", np.int32), ("res", np.int32),
    ("time", np.float32), ("dt", np.float32), ("start_time", np.float32),
    ("end_time", np.float32), ("point", np.float32, (3,)), ("normal", np.float32, (3,)),
    ("size", np.float32, (3,)), ("velocity", np.float32, (3,)), ("omega", np.float32, (3,)),
    ("rot", np.float32, (9,)), ("quat", np.float32, (4,)), ("friction", np.float32),
    ("sticky_depth", np.float32), ("contact_band", np.float32), ("band", np.float32),
    ("outer_radius", np.float32), ("inner_radius", np.float32), ("half_height", np.float32),
    ("inner_floor_z", np.float32), ("fillet_radius", np.float32),
    ("origin", np.float32, (3,)), ("cell", np.float32),
    ("sdf_val", np.uint64), ("sdf_grad", np.uint64), ("occupancy", np.uint64),
], align=True)

TICK_CTL_DTYPE = np.dtype([("box_lo", np.int32, (3,)), ("box_hi", np.int32, (3,)),
                           ("nblocks", np.int32), ("pad", np.int32)])

# every entry point declared in include/mpm_hip.h: name -> (restype, argtypes)
_P, _I32, _I64, _F = C.c_void_p, C.c_int32, C.c_int64, C.c_float
SIGNATURES = {
    "mpm_abi_version": (_I32, []),
    "mpm_status_string": (C.c_char_p, [_I32]),
    "mpm_struct_size": (_I32, [_I32]),
    "mpm_particle_pass": (_I32, [_P, _P, _P, _F, _I32, _P]),
    "mpm_grid_zero_box": (_I32, [_P, _P, _P, _I32, _P]),
    "mpm_grid_zero_blocks": (_I32, [_P, _P, _P, _I32, _I32, _P]),
    "mpm_blocks_from_box": (_I32, [_P, _P, _P, _P]),
    "mpm_grid_workspace_floats": (_I32, [_I32, _I32]),
    "mpm_grid_update": (_I32, [_P, _P, _P, _P, _P, _I32, _P, _I32, _P, _I32, _P, _F, _I32, _P]),
    "mpm_blocks_mark": (_I32, [_P, _P, _P, _P]),
    "mpm_blocks_dilate": (_I32, [_P, _P, _P, _P]),
    "mpm_blocks_compact": (_I32, [_P, _P, _P, _P, _P, _P, _P]),
    "mpm_bounds": (_I32, [_P, _P, _P]),
    "mpm_box_from_bounds": (_I32, [_P, _P, _F, _F, _P, _P]),
    "mpm_sort_keys": (_I32, [_P, _P, _P, _P, _P, _P]),
    "mpm_sort_keys_cells": (_I32, [_P, _P, _P, _P, _P, _P]),
    "mpm_sort_temp_bytes": (_I64, [_I32]),
    "mpm_sort_pairs": (_I32, [_P, _P, _P, _P, _I32, _P, _I64, _P]),
    "mpm_permute": (_I32, [_P, _P, _P, _P]),
    "mpm_sort_chunks_temp_bytes": (_I64, [_I32]),
    "mpm_sort_chunks": (_I32, [_P, _I32, _I32, _I32, _I32, _I32, _P, _P, _P, _I64, _P]),
    "mpm_pool_temp_bytes": (_I64, [_I32]),
    "mpm_pool_assign": (_I32, [_P, _P, _P, _P, _P, _P, _I64, _P]),
    "mpm_export_R": (_I32, [_P, _P, _P]),
    "mpm_export_cov": (_I32, [_P, _P, _P, _P]),
    "mpm_grid_to_dense": (_I32, [_P, _P, _P, _P, _P]),
    "mpm_cdf_zero_tags": (_I32, [_P, _P, _P, _P]),
    "mpm_cdf_copy_prev": (_I32, [_P, _P, _P, _P]),
    "mpm_cdf_stamp": (_I32, [_P, _P, _P, _P, _P]),
    "mpm_cdf_block_mask": (_I32, [_P, _P, _P, _P]),
    "mpm_cdf_set_lanes": (_I32, [_P, _P, _P]),
    "mpm_box_contact_workspace_floats": (_I32, []),
    "mpm_box_contact_wrench": (_I32, [_P, _P, _F, _I32, _P, _P, _P]),
    "mpm_rigid_workspace_bytes": (_I64, [_I32, _I32]),
    "mpm_rigid_step": (_I32, [_P, _P, _P, _I32, _P, _I32, _P, _F, _P, _P]),
    "mpm_modifier_select": (_I32, [_P, _P, _P]),
    "mpm_grid_wrench_flush": (_I32, [_P, _I32, _P, _I32, _P, _P]),
    "mpm_modifiers_apply": (_I32, [_P, _P, _I32, _P, _F, _P]),
    "mpm_mesh_field": (_I32, [_P, _I32, _P, _I32, _P, _F, _I32, _I32, _P, _P, _P]),
    "mpm_segment_distance": (_I32, [_P, _I32, _P, _F, _I32, _P, _P]),
    "mpm_grid_update_noalloc": (_I32, [_P, _P, _P, _P, _P, _I32, _P, _I32, _P, _I32, _F, _I32, _P]),
    "mpm_adjoint_g2p": (_I32, [_P, _P, _P, _P, _F, _P]),
    "mpm_adjoint_grid": (_I32, [_P, _P, _P, _P, _I32, _P, _I32, _P, _I32, _P, _F, _P]),
    "mpm_adjoint_p2g_stress": (_I32, [_P, _P, _P, _P, _F, _P]),
}

_LIB = None
ABI_VERSION = 10  # MPM_ABI_VERSION of include/mpm_hip.h this binding is written for


class NativeError(RuntimeError):
    """A C-ABI call returned a non-zero mpm_status."""


def load():
    """Load libmpm_hip.so once; raise ImportError (no fallback) when it is not built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"libmpm_hip.so not found at {LIB_PATH}: build it with "
            f"`python mpm-engine_amd/build.py` (there is no non-HIP fallback)")
    lib = C.CDLL(LIB_PATH)
    # the version first: a stale library may lack entry points bound below
    lib.mpm_abi_version.restype = _I32
    lib.mpm_abi_version.argtypes = []
    if lib.mpm_abi_version() != ABI_VERSION:
        raise ImportError(f"libmpm_hip ABI version {lib.mpm_abi_version()}, this binding needs {ABI_VERSION}: "
                          f"rebuild with `python mpm-engine_amd/build.py`")
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    sizes = {0: C.sizeof(Particles), 1: C.sizeof(Grid), 2: MODEL_DTYPE.itemsize,
             3: COLLIDER_DTYPE.itemsize, 4: TICK_CTL_DTYPE.itemsize,
             5: CDF_LANES_DTYPE.itemsize, 6: C.sizeof(CdfCollider), 7: C.sizeof(Adjoint),
             8: RIGID_BODY_DTYPE.itemsize, 9: RIGID_PLANE_DTYPE.itemsize, 10: MODIFIER_DTYPE.itemsize}
    for which, size in sizes.items():
        got = lib.mpm_struct_size(which)
        if got != size:
            raise ImportError(f"libmpm_hip ABI mismatch: struct {which} is {got} bytes in C, "
                              f"{size} in the binding")
    _LIB = lib
    return lib


def check(status: int, what: str = "") -> None:
    if status != 0:
        msg = _LIB.mpm_status_string(status).decode() if _LIB else str(status)
        raise NativeError(f"{what}: {msg} (status {status})")


def ptr(t) -> int:
    """Device (or host) address of a torch tensor, 0 for None."""
    return 0 if t is None else int(t.data_ptr())
