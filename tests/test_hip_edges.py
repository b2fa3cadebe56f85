"""Edge cases of the GPU path against the oracle: particle counts that leave ragged last
chunks (1, 255, 257, 1237 particles for 256-lane workgroups), particles left out through the
selection mask (the reference's particle_selection: nonzero skips every transfer), the
simulator-level torch import / export API (mpm_solver_warp.py:1638-1860) and the refusal of
an empty or mis-shaped particle set."""
from __future__ import annotations

import numpy as np
import pytest

from mpm_amd import Solver
from mpm_amd.core.solver import GridConfig
from oracle.sim import OracleSolver

pytestmark = pytest.mark.gpu

f32 = np.float32
GRID = GridConfig(n_grid=32, grid_lim=1.0)


def _blob(n, seed):
    rng = np.random.default_rng(seed)
    pos = rng.uniform(0.35, 0.65, (n, 3)).astype(f32)
    vel = rng.uniform(-0.3, 0.3, (n, 3)).astype(f32)
    return pos, vel


def _pair(pos, vel, **kw):
    vol = np.full(len(pos), (GRID.dx / 2) ** 3, f32)
    out = []
    for s in (Solver(grid=GRID, device="hip", **kw), OracleSolver(GRID)):
        s.load_particles(pos, vol)
        s.set_material("jelly", E=2e4, nu=0.3, density=1000.0)
        s.add_plane((0, 0, 3 * GRID.dx), (0, 0, 1), "sticky")
        s.set_v(vel)
        out.append(s)
    return out


@pytest.mark.parametrize("n", [1, 255, 257, 1237])
def test_ragged_particle_counts_match_oracle(n):
    pos, vel = _blob(n, n)
    hip, ora = _pair(pos, vel)
    for _ in range(2):
        hip.step(1e-4, 10)
        ora.step(1e-4, 10)
    err = np.abs(hip.x() - ora.x()).max()
    print("n", n, "max |x - x_oracle| / dx", err / GRID.dx)
    assert err <= 1e-4 * GRID.dx
    np.testing.assert_allclose(hip.v(), ora.v(), atol=1e-4 * max(np.abs(ora.v()).max(), 1e-3))


@pytest.mark.parametrize("deterministic", [None, False])
def test_selection_leaves_particles_out(deterministic):
    import torch

    pos, vel = _blob(3000, 7)
    hip, ora = _pair(pos, vel, deterministic=deterministic)
    sel = (np.random.default_rng(8).uniform(0, 1, len(pos)) < 0.2).astype(np.int32)
    hip._sim.import_particle_selection_from_torch(torch.from_numpy(sel))
    ora._sim.selection[:] = sel
    np.testing.assert_array_equal(hip._sim.export_particle_selection_to_torch().cpu().numpy(), sel)
    for _ in range(2):
        hip.step(1e-4, 10)
        ora.step(1e-4, 10)
    out = sel != 0
    # left-out particles keep their state bit for bit; the others match the oracle
    np.testing.assert_array_equal(hip.x()[out], pos[out])
    np.testing.assert_array_equal(hip.v()[out], vel[out])
    assert np.abs(hip.x() - ora.x()).max() <= 1e-4 * GRID.dx
    assert np.abs(hip.x()[~out] - pos[~out]).max() > 1e-4  # the rest did move


def test_simulator_torch_import_export_round_trip():
    import torch

    pos, vel = _blob(500, 3)
    hip, _ = _pair(pos, vel)
    sim = hip._sim
    n = len(pos)
    rng = np.random.default_rng(4)
    F = (np.eye(3) + 0.05 * rng.standard_normal((n, 3, 3))).astype(f32)
    C = rng.standard_normal((n, 3, 3)).astype(f32)
    x2 = (pos + 0.01).astype(f32)
    sim.import_particle_x_from_torch(torch.from_numpy(x2))
    sim.import_particle_v_from_torch(torch.from_numpy(-vel))
    sim.import_particle_F_from_torch(torch.from_numpy(F))
    sim.import_particle_C_from_torch(torch.from_numpy(C.reshape(n, 9)))
    np.testing.assert_array_equal(sim.export_particle_x_to_torch().cpu().numpy(), x2)
    np.testing.assert_array_equal(sim.export_particle_v_to_torch().cpu().numpy(), -vel)
    np.testing.assert_array_equal(sim.export_particle_F_to_torch().cpu().numpy(), F.reshape(n, 9))
    np.testing.assert_array_equal(sim.export_particle_C_to_torch().cpu().numpy(), C.reshape(n, 9))
    assert sim.export_particle_vol_to_torch().shape == (n,)
    assert sim.export_particle_stress_to_torch().shape == (n, 9)
    assert sim.export_particle_L_to_torch().shape == (n, 9)
    assert sim.export_particle_cov_to_torch().shape == (6 * n,)
    # R^T of the polar factor of F_trial (compute_R_from_F reads F_trial), (n, 9) row-major
    sim.F_trial.copy_(torch.from_numpy(F.reshape(n, 9).T.copy()).to(sim.device))
    R = sim.export_particle_R_to_torch().cpu().numpy().reshape(n, 3, 3).transpose(0, 2, 1)
    U, _, Vt = np.linalg.svd(F.astype(np.float64))
    np.testing.assert_allclose(R, U @ Vt, atol=2e-5)
    m = sim.export_grid_m_to_torch()
    assert tuple(m.shape) == (32, 32, 32)
    assert tuple(sim.export_grid_v_out_to_torch().shape) == (32, 32, 32, 3)
    with pytest.raises(ValueError, match="rows"):
        sim.import_particle_x_from_torch(torch.zeros(n + 1, 3))
    # reset_densities_and_update_masses: mass = density * vol, and the deterministic
    # scatter's mass scale follows the new heaviest particle
    dens = torch.from_numpy(np.random.default_rng(6).uniform(500.0, 3000.0, n).astype(np.float32))
    sim.reset_densities_and_update_masses(dens)
    np.testing.assert_allclose(sim.mass.cpu().numpy(), dens.numpy() * sim.vol.cpu().numpy(), rtol=1e-6)
    np.testing.assert_array_equal(sim.density.cpu().numpy(), dens.numpy())  # particle_density kept
    assert sim._m_max is None
    hip.step(1e-4, 2)
    assert sim._m_max == pytest.approx(float((dens * sim.vol.cpu()).max()), rel=1e-6)
    assert np.isfinite(hip.x()).all()
    # material import re-selects the pass specialisation (here: all viscous)
    sim.import_particle_material_from_torch(torch.full((n,), 6, dtype=torch.int32))
    assert sim._mat_ids == {6}
    np.testing.assert_array_equal(sim.export_particle_material_to_torch().cpu().numpy(), np.full(n, 6))


def test_empty_or_misshaped_particle_sets_are_refused():
    s = Solver(grid=GRID, device="hip")
    with pytest.raises(ValueError, match="non-empty"):
        s.load_particles(np.zeros((0, 3), f32), np.zeros(0, f32))
    with pytest.raises(ValueError, match="non-empty"):
        s.load_particles(np.zeros((10, 2), f32), np.zeros(10, f32))
    with pytest.raises(ValueError, match="volumes"):
        s.load_particles(np.full((10, 3), 0.5, f32), np.zeros(9, f32))
