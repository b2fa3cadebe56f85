"""The oracle's tick orchestration restates the reference's schedules; these are the
reference's own pipeline-equivalence and analytic gates, run on the oracle (CPU):
fused == split bitwise (tests/test_fused_pipeline.py:59-75), restricted == full launches
(tests/test_restricted_launch.py:34-48), sparse == dense (tests/test_sparse_blocks.py:
37-48), exact-multiset block sort (tests/test_particle_sort.py:38-46), free-fall momentum
and volume (tests/test_engine.py:51-73) and the grid-edge guard (test_restricted_launch.py:
68-79)."""
from __future__ import annotations

import numpy as np
import pytest

from mpm_amd import GlassProfile, GridConfig
from mpm_amd.geometry import sphere_sdf
from mpm_amd.materials import elastic, newtonian, vonmises
from oracle.sim import OracleSolver

PROF = GlassProfile(outer_radius=0.05, inner_radius=0.04, height=0.12, base_thickness=0.02,
                    fillet_radius=0.006)


def _blob(n=1200, seed=0):
    rng = np.random.default_rng(seed)
    pts = (rng.random((n, 3), dtype=np.float32) * 0.10 + 0.15).astype(np.float32)
    return GridConfig(n_grid=40, grid_lim=0.4), pts, np.full(n, 1.0e-7, np.float32)


def _run(material, fused, ticks=4, substeps=6, cup=False, **kw):
    grid, pts, vol = _blob()
    s = OracleSolver(grid, fused=fused, **kw).load_particles(pts, vol)
    s.set_material(material)
    s.add_plane((0, 0, 3 * grid.dx), (0, 0, 1), "separate", friction=0.3)
    s.add_domain_walls()
    h = None
    if cup:
        c = np.array([0.2, 0.2, 0.2])  # the glass wall cuts through the blob
        h = s.add_cup(PROF, c)
        s.set_cup(h, center=c, quat=(1.0, 0, 0, 0), velocity=(0.02, 0, 0), omega=(0, 0, 0.5))
    w = None
    for _ in range(ticks):
        if h is not None:
            s.reset_cup_wrench(h)
        s.step(2.0e-4, substeps)
        if h is not None:
            w = s.cup_wrench(h, 2.0e-4 * substeps)
    return s, w


MATERIALS = [("newtonian", newtonian(eta=2.0, density=1000.0, bulk_modulus=2.0e5)),
             ("jelly", elastic(E=5.0e4, nu=0.3, density=800.0)),
             ("metal", vonmises(E=1.0e5, nu=0.3, yield_stress=2.0e3, density=1200.0))]


@pytest.mark.parametrize("label,material", MATERIALS, ids=[m[0] for m in MATERIALS])
def test_oracle_fused_bitwise_equals_split(label, material):
    a, _ = _run(material, fused=False)
    b, _ = _run(material, fused=True)
    for f in ("x", "v", "F", "stress"):
        assert np.array_equal(getattr(a, f)(), getattr(b, f)()), f"{label}: {f} diverged"


def test_oracle_fused_bitwise_with_moving_cup_and_wrench():
    a, wa = _run(MATERIALS[0][1], fused=False, cup=True)
    b, wb = _run(MATERIALS[0][1], fused=True, cup=True)
    assert np.array_equal(a.x(), b.x())
    np.testing.assert_array_equal(wa["force"], wb["force"])
    np.testing.assert_array_equal(wa["torque"], wb["torque"])
    assert np.abs(wa["force"]).sum() > 0


def test_oracle_restricted_equals_full():
    def scene(restrict):
        grid, pts, vol = _blob(seed=2)
        s = OracleSolver(grid).load_particles(pts, vol)
        s.set_material(newtonian(eta=5.0, density=1000.0, bulk_modulus=5.0e5))
        s.add_plane((0, 0, 3 * grid.dx), (0, 0, 1), "slip", friction=0.2)
        h = s.add_sdf_collider(sphere_sdf(0.03, res=24), center=(0.2, 0.2, 0.09))
        s.add_box((0.30, 0.20, 0.10), (0.02, 0.02, 0.02), velocity=(-0.05, 0.0, -0.02))
        s._sim.restrict_bc = restrict
        s._sim.restrict_grid = restrict
        return s, h

    sa, ha = scene(True)
    sb, hb = scene(False)
    for _ in range(8):
        sa.reset_sdf_force(ha)
        sb.reset_sdf_force(hb)
        sa.step(2e-4, 4)
        sb.step(2e-4, 4)
    assert np.array_equal(sa.x(), sb.x())
    np.testing.assert_allclose(sa.sdf_wrench(ha, 8e-4)["force"], sb.sdf_wrench(hb, 8e-4)["force"],
                               rtol=1e-4, atol=1e-6)


def test_oracle_sparse_equals_dense():
    def scene(sparse):
        grid = GridConfig(n_grid=48, grid_lim=0.4)
        h = grid.dx / 2
        pts = []
        for cx in (0.08, 0.32):
            ax = np.arange(-0.015, 0.015, h)
            g = np.stack(np.meshgrid(ax, ax, np.arange(0, 0.03, h), indexing="ij"), -1).reshape(-1, 3)
            pts.append(g + [cx, 0.2, 0.06])
        pts = np.concatenate(pts).astype(np.float32)
        s = OracleSolver(grid, sparse=sparse).load_particles(pts, np.full(len(pts), h ** 3, np.float32))
        s.set_material(newtonian(eta=5.0, density=1000.0, bulk_modulus=5.0e5))
        s.add_plane((0, 0, 3 * grid.dx), (0, 0, 1), "slip", friction=0.2)
        return s

    a, b = scene(True), scene(False)
    for _ in range(6):
        a.step(2e-4, 3)
        b.step(2e-4, 3)
    assert np.array_equal(a.x(), b.x())
    assert 0.0 < a._sim.active_block_fraction() < 0.15


def test_oracle_sort_is_exact_multiset():
    grid, pts, vol = _blob(n=800, seed=4)
    rng = np.random.default_rng(1)
    pts = pts[rng.permutation(len(pts))]
    s = OracleSolver(grid).load_particles(pts, vol)
    s.set_material(newtonian(eta=2.0, density=1000.0, bulk_modulus=2.0e5))
    s.step(2e-4, 3)
    rows = lambda: np.concatenate([s.x(), s.v(), s.F().reshape(len(pts), -1)], 1)  # noqa: E731
    before = rows()
    assert s._sort_particles() is True
    after = rows()
    assert np.array_equal(before[np.lexsort(before.T)], after[np.lexsort(after.T)])
    assert s._sort_particles() is False


def test_oracle_free_fall_momentum_and_volume():
    grid = GridConfig(n_grid=40, grid_lim=0.4)
    h = grid.dx / 2
    ax = np.arange(-0.02, 0.02, h)
    pts = (np.stack(np.meshgrid(ax, ax, ax, indexing="ij"), -1).reshape(-1, 3) + 0.2).astype(np.float32)
    s = OracleSolver(grid).load_particles(pts, np.full(len(pts), h ** 3, np.float32))
    s.set_material(newtonian(eta=40.0, density=1000.0).with_yield(200.0))
    dt, n = 2e-5, 100
    s.step(dt, n)
    vz = float(s.v()[:, 2].mean())
    assert abs(vz + 9.81 * dt * n) < 0.05 * 9.81 * dt * n
    J = np.abs(np.linalg.det(s.F())).mean()
    assert abs(J - 1.0) < 5e-3


def test_oracle_grid_edge_raises():
    grid = GridConfig(n_grid=32, grid_lim=0.4)
    s = OracleSolver(grid).load_particles(np.array([[0.2, 0.2, 0.4 - 2 * grid.dx]], np.float32),
                                          np.full(1, 1e-6, np.float32))
    s.set_material(newtonian(eta=5.0))
    with pytest.raises(RuntimeError, match="out of bounds"):
        s.step(2e-4, 1)
