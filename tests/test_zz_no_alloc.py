"""mpm_grid_update_noalloc (the allocation-free grid update, scratch in a device array of
the library) against mpm_grid_update with a caller workspace, through the C ABI (GPU).

Both run on the same grid state: a scene's accumulators after one scatter and the
collider rows of its last tick. Their results (node velocities, cleared accumulators,
block states, wrench) must be bitwise equal: same kernels, same fixed-order reduction.
Two calls in a row, the second on a smaller scene (fewer blocks, fewer wrench slots),
for the fixed-point (deterministic) and the float scatter; then more than 64 wrench
slots (the integer-atomic row that must start at zero), twice; then the refusals."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from mpm_amd import GlassProfile, GridConfig, Solver
from mpm_amd.core import native
from mpm_amd.geometry import sphere_sdf
from mpm_amd.materials import newtonian

pytestmark = pytest.mark.gpu

DT = 2.0e-4
PROF = GlassProfile(outer_radius=0.05, inner_radius=0.04, height=0.12, base_thickness=0.02,
                    fillet_radius=0.006)


def _scene(n_grid, n, deterministic, big):
    rng = np.random.default_rng(n_grid + n)
    grid = GridConfig(n_grid=n_grid, grid_lim=0.4)
    pts = (rng.random((n, 3)) * 0.12 + np.array([0.14, 0.14, 0.12])).astype(np.float32)
    s = Solver(grid=grid, device="hip", deterministic=deterministic)
    s.load_particles(pts, np.full(n, 1e-7, np.float32))
    s.set_material(newtonian(eta=2.0, density=1000.0, bulk_modulus=2.0e5))
    s.add_plane((0, 0, 3 * grid.dx), (0, 0, 1), "separate", friction=0.3)
    s.add_box((0.2, 0.2, 0.26), (0.05, 0.05, 0.02), velocity=(0.0, 0.0, -0.05))
    if big:
        s.add_cup(PROF, (0.2, 0.2, 0.2), (1.0, 0.0, 0.0, 0.0), omega=(0.0, 0.5, 0.0), friction=0.1)
        s.add_sdf_collider(sphere_sdf(0.03, res=24), center=(0.2, 0.25, 0.2), omega=(0.0, 0.0, 2.0))
    s.add_domain_walls()
    v = np.zeros((n, 3), np.float32)
    v[:, 2] = -0.5
    s.set_v(v)
    s.step(DT, 4)  # develops the state; leaves the tick's collider rows, launch box and block list
    return s


def _prepare(s):
    """One fresh scatter into the scene's accumulators; returns the grid state the grid
    update starts from (accumulators, node velocities, block states)."""
    sim = s._sim
    if sim.deterministic:
        sim._fix_scales(DT)
    sim._pass(native.PASS_STRESS | native.PASS_P2G, DT, sim.stream())
    acc = sim.node_fix if sim.deterministic else sim.node_in
    return acc.clone(), sim.node_out.clone(), None if sim.block_touch is None else sim.block_touch.clone()


def _call(s, snap, noalloc, slots=None):
    """One grid update from the state `snap` by either entry point; returns (node_out,
    accumulators after, block states, wrench)."""
    import torch

    sim = s._sim
    lib = sim.lib
    st = sim.stream()
    det = sim.deterministic
    acc = sim.node_fix if det else sim.node_in
    acc.copy_(snap[0])
    sim.node_out.copy_(snap[1])
    if snap[2] is not None:
        sim.block_touch.copy_(snap[2])
    slots = sim.n_slots if slots is None else slots
    K = len(sim.colliders.items)
    flags = native.GRID_FIXED_IN if det else native.GRID_ZERO_IN
    wrench = torch.zeros(6 * max(slots, 1), dtype=torch.float32, device=sim.device)
    args = (sim.pG, native.ptr(sim.model_dev), native.ptr(sim.ctl), native.ptr(sim.box_list),
            native.ptr(sim.ctl[6:]), sim.n_blocks, native.ptr(sim.sched_dev) if K else None, K,
            native.ptr(wrench), slots)
    if noalloc:
        rc = lib.mpm_grid_update_noalloc(*args, DT, flags, st)
    else:
        ws = torch.zeros(max(int(lib.mpm_grid_workspace_floats(sim.n_blocks, slots)), 1), dtype=torch.float32,
                         device=sim.device)
        rc = lib.mpm_grid_update(*args, native.ptr(ws), DT, flags, st)
    assert rc == 0, lib.mpm_status_string(rc)
    torch.cuda.synchronize()
    return (sim.node_out.clone(), acc.clone(), None if sim.block_touch is None else sim.block_touch.clone(),
            wrench.clone())


def _same(a, b):
    for x, y in zip(a, b):
        if x is None:
            assert y is None
            continue
        assert torch_equal(x, y)


def torch_equal(x, y):
    import torch

    return bool(torch.equal(x.view(torch.int32) if x.dtype == torch.float32 else x,
                            y.view(torch.int32) if y.dtype == torch.float32 else y))


@pytest.mark.parametrize("deterministic", [True, False], ids=["fixed_point", "float"])
def test_noalloc_grid_update_equals_workspace_entry_point(deterministic):
    for n_grid, n, big in ((64, 20000, True), (32, 3000, False)):  # second call: smaller input
        s = _scene(n_grid, n, deterministic, big)
        snap = _prepare(s)
        ref = _call(s, snap, noalloc=False)
        got = _call(s, snap, noalloc=True)
        _same(ref, got)
        assert float(ref[3].abs().max()) > 0.0  # the wrench saw contacts
        assert not torch_equal(ref[0], snap[1])  # the update did something


def test_noalloc_above_64_slots_starts_from_a_clear_row():
    """More than 64 wrench slots: the fixed-point wrench is summed with integer atomics
    into one row of the shared array, which the entry point clears first; run twice so
    the second call finds the first call's array."""
    for n_grid, n, big in ((64, 20000, True), (32, 3000, False)):
        s = _scene(n_grid, n, True, big)
        snap = _prepare(s)
        ref = _call(s, snap, noalloc=False, slots=70)
        got = _call(s, snap, noalloc=True, slots=70)
        _same(ref, got)
        got2 = _call(s, snap, noalloc=True, slots=70)
        _same(ref, got2)


def test_noalloc_refusals():
    s = _scene(32, 3000, True, False)
    sim = s._sim
    lib = sim.lib
    lists = (native.ptr(sim.box_list), native.ptr(sim.ctl[6:]))
    base = (sim.pG, native.ptr(sim.model_dev), native.ptr(sim.ctl), lists[0], lists[1], sim.n_blocks,
            native.ptr(sim.sched_dev), len(sim.colliders.items), native.ptr(sim.wrench))
    # beyond the array: 12 * 70000 floats of integer-atomic row
    assert lib.mpm_grid_update_noalloc(*base, 70000, C.c_float(DT), native.GRID_FIXED_IN, sim.stream()) == 3
    # deferred rows cannot outlive a call in the shared array
    assert lib.mpm_grid_update_noalloc(*base, sim.n_slots, C.c_float(DT),
                                       native.GRID_FIXED_IN | native.GRID_WRENCH_DEFER, sim.stream()) == 1
    assert b"documented maximum" in lib.mpm_status_string(3)
