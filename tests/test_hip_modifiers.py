"""Particle modifiers (pre-P2G impulses, particle velocity translation and rotation) on the
GPU against the oracle (mpm_modifier_select / mpm_modifiers_apply; the reference's
add_impulse_on_particles and enforce_particle_velocity_*, mpm_solver_warp.py:2385-2539,
applied by the split substep, :1181-1196).

Scene: a jelly slab with a translated box, a rotating cylinder and an impulse region
whose window closes mid-tick, 4 ticks of 10 substeps, with and without a block sort
every tick (the selection masks must follow the particles)."""
from __future__ import annotations

import numpy as np
import pytest

from mpm_amd import Solver
from mpm_amd.core.solver import GridConfig
from oracle.sim import OracleSolver

pytestmark = pytest.mark.gpu

f32 = np.float32


def _scene(cls, sort_interval=0, **kw):
    g = GridConfig(n_grid=32, grid_lim=1.0)
    h = g.dx / 2
    ax = np.arange(0.3 + h / 2, 0.7, h)
    pts = np.stack(np.meshgrid(ax, ax, np.arange(0.3 + h / 2, 0.5, h), indexing="ij"), -1).reshape(-1, 3).astype(f32)
    s = cls(grid=g, sort_interval=sort_interval, **kw) if cls is Solver else cls(g, sort_interval=sort_interval)
    s.load_particles(pts, np.full(len(pts), h ** 3, f32))
    s.set_material("jelly", E=1e5, nu=0.3, density=1000.0)
    s.add_plane((0, 0, 3 * g.dx), (0, 0, 1), "sticky")
    sim = s._sim
    sim.enforce_particle_velocity_translation((0.35, 0.5, 0.4), (0.04, 0.3, 0.3), (0.2, 0.0, 0.05), 0.0, 2.5e-3)
    sim.enforce_particle_velocity_rotation((0.6, 0.5, 0.4), (0.2, 0.1, 1.0), (0.3, 0.08), 3.0, 0.1, 0.0, 1.0)
    sim.add_impulse_on_particles((0.0, 0.0, 2e-3), 1e-4, point=(0.5, 0.35, 0.4), size=(0.05, 0.04, 0.3),
                                 num_dt=15)
    return s


@pytest.mark.parametrize("sort_interval", [0, 1])
def test_modifiers_match_oracle(sort_interval):
    hip = _scene(Solver, sort_interval=sort_interval, device="hip")
    ora = _scene(OracleSolver, sort_interval=sort_interval)
    for _ in range(4):
        hip.step(1e-4, 10)
        ora.step(1e-4, 10)
    dx = 1.0 / 32
    err = np.abs(hip.x() - ora.x()).max()
    print("max |x - x_oracle| / dx", err / dx, "max |v|", np.abs(ora.v()).max())
    assert np.abs(ora.v()).max() > 0.1
    assert err <= 1e-4 * dx
    np.testing.assert_allclose(hip.v(), ora.v(), atol=1e-4 * np.abs(ora.v()).max())


def test_modifier_count_limit():
    s = _scene(Solver, device="hip")
    for _ in range(29):
        s._sim.add_impulse_on_particles((0, 0, 1.0), 1e-4)
    with pytest.raises(ValueError, match="at most 32"):
        s._sim.add_impulse_on_particles((0, 0, 1.0), 1e-4)
