"""The C-ABI library loads without a GPU and exports every entry point declared in
include/mpm_hip.h, with struct layouts matching the Python binding (no compute calls)."""
from __future__ import annotations

import ctypes
import os
import re

from mpm_amd.core import native

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include",
                      "mpm_hip.h")


def _declared():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(mpm_[A-Za-z_0-9]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    lib = native.load()
    names = _declared()
    assert len(names) >= 15
    for nm in names:
        assert hasattr(lib, nm), f"{nm} declared in include/mpm_hip.h but not exported"
    assert set(names) == set(native.SIGNATURES), "binding and header disagree"


def test_struct_layouts_and_introspection():
    lib = native.load()
    assert lib.mpm_abi_version() == native.ABI_VERSION
    assert lib.mpm_struct_size(1) == ctypes.sizeof(native.Grid)
    assert lib.mpm_struct_size(2) == native.MODEL_DTYPE.itemsize
    assert lib.mpm_struct_size(3) == native.COLLIDER_DTYPE.itemsize
    assert lib.mpm_struct_size(0) == ctypes.sizeof(native.Particles)
    assert lib.mpm_struct_size(5) == native.CDF_LANES_DTYPE.itemsize
    assert lib.mpm_struct_size(6) == ctypes.sizeof(native.CdfCollider)
    assert lib.mpm_struct_size(7) == ctypes.sizeof(native.Adjoint)
    assert lib.mpm_struct_size(8) == native.RIGID_BODY_DTYPE.itemsize
    assert lib.mpm_struct_size(9) == native.RIGID_PLANE_DTYPE.itemsize
    assert lib.mpm_struct_size(10) == native.MODIFIER_DTYPE.itemsize
    assert lib.mpm_struct_size(11) == -1
    assert lib.mpm_status_string(1) == b"invalid argument"


def test_entry_points_reject_null_arguments_without_touching_the_gpu():
    lib = native.load()
    assert lib.mpm_particle_pass(None, None, None, 1e-4, 7, None) == 1
    assert lib.mpm_grid_update(None, None, None, None, None, 0, None, 0, None, 0, None, 1e-4, 1, None) == 1
    assert lib.mpm_grid_workspace_floats(100, -1) == -1
    assert lib.mpm_grid_workspace_floats(100, 2) == 25 * 2 * 12  # fixed-order partials (float or int64)
    assert lib.mpm_grid_workspace_floats(100, 65) == 12 * 65     # atomic path: fixed-point wrench only
    assert lib.mpm_permute(None, None, None, None) == 1
