"""Rigid bodies (material 8) on the GPU against the oracle: the split substep's rigid step
(mpm_rigid_step: momentum gather, body integration, restitution planes, particle update;
the reference's _p2g2p_tail, mpm_solver_warp.py:1338-1367, and finalize_rigid_bodies,
:834-873).

Scenes: a tilted rigid cube (the lattice rotated in the world, so the deepest particle at
impact is one corner, not a tie along a face) falling onto a restitution plane next to a
jelly block, and the same with a block sort every tick. Checked: body state (COM,
velocity, angular velocity, orientation) and particle positions against the oracle after
the impact; the orientation stays orthonormal; the API's refusals."""
from __future__ import annotations

import numpy as np
import pytest

from mpm_amd import Solver
from mpm_amd.core.solver import GridConfig
from mpm_amd.materials import elastic
from oracle.sim import OracleSolver

pytestmark = pytest.mark.gpu

f32 = np.float32


def _scene(cls, sort_interval=0, restitution=0.6, friction=0.3, **kw):
    g = GridConfig(n_grid=32, grid_lim=1.0)
    h = g.dx / 2
    ax = np.arange(-0.08 + h / 2, 0.08, h)
    cube = np.stack(np.meshgrid(ax, ax, ax, indexing="ij"), -1).reshape(-1, 3)
    a, b = 0.4, 0.3
    Rx = np.array([[1, 0, 0], [0, np.cos(a), -np.sin(a)], [0, np.sin(a), np.cos(a)]])
    Ry = np.array([[np.cos(b), 0, np.sin(b)], [0, 1, 0], [-np.sin(b), 0, np.cos(b)]])
    cube = cube @ (Ry @ Rx).T
    floor = 3 * g.dx
    cube = (cube + [0.35, 0.5, floor + 0.02 - cube[:, 2].min()]).astype(f32)  # corner 2 cm up
    jelly = (np.stack(np.meshgrid(np.arange(0.56, 0.72, h), np.arange(0.42, 0.58, h), np.arange(0.12, 0.24, h),
                                  indexing="ij"), -1).reshape(-1, 3)).astype(f32)
    pts = np.concatenate([cube, jelly])
    vol = np.full(len(pts), h ** 3, f32)
    s = cls(grid=g, sort_interval=sort_interval, **kw) if cls is Solver else cls(g, sort_interval=sort_interval)
    s.load_particles(pts, vol)
    s.set_material(elastic(E=2e5, nu=0.3, density=1000.0))
    s.set_material_range(0, len(cube), "rigid", obj_id=0, density=800.0)
    s.add_plane((0, 0, 3 * g.dx), (0, 0, 1), "separate", friction=friction, restitution=restitution)
    s.add_domain_walls()
    s.finalize_rigid_bodies()
    v = np.zeros((len(pts), 3), f32)
    v[:len(cube)] = (0.3, -0.1, -1.5)  # the body takes its velocity from its particles' grid momentum
    if cls is Solver:
        s.set_v(v)
    else:
        s._sim.v[:] = v
    return s, len(cube)


def _run(s, ticks=8, substeps=20, dt=1e-4):
    for _ in range(ticks):
        s.step(dt, substeps)
    return s


def _compare(hip, ora, n_cube):
    h, o = hip.rigid_state(0), ora.rigid_state(0)
    print({k: (h[k], o[k]) for k in h})
    np.testing.assert_allclose(h["com"], o["com"], atol=1e-5)
    np.testing.assert_allclose(h["v"], o["v"], atol=2e-4 * max(1.0, np.abs(o["v"]).max()))
    np.testing.assert_allclose(h["omega"], o["omega"], atol=2e-3 * max(1.0, np.abs(o["omega"]).max()))
    np.testing.assert_allclose(h["R"], o["R"], atol=2e-5)
    np.testing.assert_allclose(h["R"] @ h["R"].T, np.eye(3), atol=2e-6)
    dx = 1.0 / 32
    err = np.abs(hip.x() - ora.x()).max()
    print("max |x - x_oracle| / dx", err / dx)
    assert err <= 1e-3 * dx


def test_rigid_body_impact_matches_oracle():
    """160 substeps: the cube, thrown down at 1.5 m/s from 2 cm, hits the restitution plane
    (e = 0.6, mu = 0.3) at a corner and rebounds spinning, beside a jelly block falling
    onto the same floor."""
    hip, n_cube = _scene(Solver, device="hip")
    ora, _ = _scene(OracleSolver)
    _run(hip)
    _run(ora)
    o = ora.rigid_state(0)
    assert np.abs(o["omega"]).max() > 0.1  # the corner impact set it spinning
    _compare(hip, ora, n_cube)


def test_rigid_body_with_block_sort_matches_oracle():
    """The same scene with a block sort before every tick: rigid ids and body-frame
    reference positions move with the particles."""
    hip, n_cube = _scene(Solver, device="hip", sort_interval=1)
    ora, _ = _scene(OracleSolver, sort_interval=1)
    _run(hip, ticks=4)
    _run(ora, ticks=4)
    _compare(hip, ora, n_cube)


def test_rigid_api_refusals():
    g = GridConfig(n_grid=16, grid_lim=1.0)
    pts = np.full((8, 3), 0.5, f32) + np.arange(8, dtype=f32)[:, None] * 0.01
    s = Solver(grid=g, device="hip", deterministic=True)
    s.load_particles(pts, np.full(8, 1e-6, f32))
    s.set_material("rigid", density=1000.0)
    with pytest.raises(NotImplementedError, match="deterministic"):
        s.finalize_rigid_bodies()
    with pytest.raises(ValueError, match="restitution"):
        s.add_plane((0, 0, 0.2), (0, 0, 1), "separate", restitution=1.5)
    t = Solver(grid=g, device="hip")
    t.load_particles(pts, np.full(8, 1e-6, f32))
    t.set_material("rigid", density=1000.0)
    t.finalize_rigid_bodies()
    with pytest.raises(ValueError, match="no rigid body"):
        t.rigid_state(3)
