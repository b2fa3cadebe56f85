"""Batched scenes (BatchSolver) vs the CPU oracle run scene by scene (GPU).

Every scene of a batch must evolve exactly as that scene alone: positions after 50
substeps within 1e-4 dx of the oracle's single-scene run and box reaction forces within
1e-3 relative, whatever the launch shape, sort interval, sparse / dense mode or graph
replay. Float atomics accumulate in a different order than the oracle's sequential loop,
which is the only expected source of difference."""
from __future__ import annotations

import numpy as np
import pytest

from mpm_amd import GridConfig
from mpm_amd.batch import LAUNCH_SHAPES, BatchSolver
from mpm_amd.scenes import DoughPressBatch, herschel_bulkley_dough
from oracle.sim import OracleSolver

pytestmark = pytest.mark.gpu

DT = 5e-5


def _scenes(n_scenes, seed=1):
    return DoughPressBatch(n_scenes=n_scenes, n_grid=32, n_min=300, n_max=1500, seed=seed, dt=DT)


def _batch(sc, **kw):
    b = BatchSolver(grid=sc.grid, device="hip", **kw)
    return sc.build_batch(b)


def _oracle(sc, s):
    o = OracleSolver(sc.grid)
    o.load_particles(sc.pos[s], sc.vol[s])
    o.set_material(herschel_bulkley_dough())
    o._sim.yield_stress[:] = sc.yield_stress[s]
    o.add_plane((0, 0, sc.floor), (0, 0, 1), "sticky")
    o.add_domain_walls()
    h = o.add_box(sc.box_center[s], sc.box_half[s], (0.0, 0.0, -sc.box_speed[s]))
    return o, h


def _run(solver, ticks=2, substeps=25):
    for _ in range(ticks):
        solver.step(DT, substeps)
    return solver


def test_batch_matches_oracle_scene_by_scene():
    """36 scenes with 2 wrench slots each (72 > 64: the atomic wrench path), 2 ticks of
    25 substeps, sparse blocks and block sort on: each scene's positions and tool force
    against that scene run alone through the oracle."""
    sc = _scenes(36)
    b = _run(_batch(sc))
    T = 2 * 25 * DT
    dx = sc.grid.dx
    worst = 0.0
    for s in range(36):
        o, h = _oracle(sc, s)
        _run(o)
        err = np.abs(b.x(s) - o.x()).max()
        worst = max(worst, err)
        assert err <= 1e-4 * dx, f"scene {s}: max |dx| {err:.3e} > 1e-4 dx"
        f_b = b.tool_force(s, sc.tool[s], T)
        f_o = o.tool_force(h, T)
        scale = np.abs(f_o).max()
        assert np.allclose(f_b, f_o, rtol=1e-3, atol=1e-3 * scale + 1e-9), (s, f_b, f_o)
    print(f"worst position error {worst / dx:.2e} dx")


def test_batch_fixed_order_wrench_path():
    """Few scenes (wrench slots <= 64: fixed-order reduction) agree with the oracle too."""
    sc = _scenes(3, seed=2)
    b = _run(_batch(sc))
    T = 2 * 25 * DT
    for s in range(3):
        o, h = _oracle(sc, s)
        _run(o)
        assert np.abs(b.x(s) - o.x()).max() <= 1e-4 * sc.grid.dx
        f_o = o.tool_force(h, T)
        assert np.allclose(b.tool_force(s, sc.tool[s], T), f_o, rtol=1e-3,
                           atol=1e-3 * np.abs(f_o).max() + 1e-9)


@pytest.mark.parametrize("shape", LAUNCH_SHAPES)
def test_launch_shapes_agree(shape):
    """Every particle-pass launch shape the autotuner may pick gives the same rollout."""
    sc = _scenes(6, seed=3)
    ref = _run(_batch(sc))
    b = _batch(sc)
    b.set_launch(*shape)
    _run(b)
    for s in range(6):
        assert np.abs(b.x(s) - ref.x(s)).max() <= 1e-4 * sc.grid.dx


def test_paired_pass_is_bit_identical_in_the_block_pool():
    """The paired interior pass over the shared block pool (pool slots through node_addr,
    per-scene tiles) against the single-particle pass: the fixed-point sums make every
    position bitwise equal."""
    sc = _scenes(6, seed=3)
    a, b = _batch(sc), _batch(sc)
    a.set_launch(256, 1, True)
    _run(a)
    _run(b)
    assert a._sim.pool and a._sim.chunk_count.tolist()[1] > 0
    for s in range(6):
        assert np.array_equal(a.x(s), b.x(s))


def test_dense_and_sparse_and_unsorted_agree():
    sc = _scenes(5, seed=4)
    ref = _run(_batch(sc))
    for kw in ({"sparse": False}, {"sort_interval": 0}, {"sparse": False, "sort_interval": 3}):
        b = _run(_batch(sc, **kw))
        for s in range(5):
            assert np.abs(b.x(s) - ref.x(s)).max() <= 1e-4 * sc.grid.dx, kw


def test_block_pool_matches_dense_storage_bitwise():
    """The shared block pool (the default with sparse stepping) against dense per-scene
    storage under the same stepping: a block's fixed-point sums do not depend on where it
    is stored, so positions, velocities, box forces and the dense grid readout agree bit
    for bit. The pool starts at 64 blocks and grows at the first rebuild (5 scenes touch
    several hundred), and holds the grid in a fraction of the dense bytes."""
    import torch

    sc = _scenes(5, seed=4)
    a = _run(_batch(sc))
    b = _run(_batch(sc, pool=False))
    assert a._sim.pool and not b._sim.pool
    active = int(a._sim._blk["counts"][0].item())
    assert 64 < active <= a._sim.pool_blocks < a._sim.n_blocks
    T = 2 * 25 * DT
    for s in range(5):
        assert np.array_equal(a.x(s), b.x(s)), s
        assert np.array_equal(a.v(s), b.v(s)), s
        assert np.array_equal(a.tool_force(s, sc.tool[s], T), b.tool_force(s, sc.tool[s], T)), s
    for ga, gb in zip(a.grid_dense(), b.grid_dense()):
        assert torch.equal(ga, gb)
    print(f"grid bytes: pool {a._sim.grid_bytes()}, dense {b._sim.grid_bytes()} ({active} active of "
          f"{a._sim.n_blocks} blocks)")
    assert a._sim.grid_bytes() < 0.5 * b._sim.grid_bytes()


def test_float_and_fixed_point_batches_agree():
    """The float-atomic batch and the default fixed-point one follow each other within the
    parity bar, and the fixed-point batch repeats bit for bit."""
    sc = _scenes(6, seed=10)
    a = _run(_batch(sc))
    b = _run(_batch(sc))
    c = _run(_batch(sc, deterministic=False))
    for s in range(6):
        assert np.array_equal(a.x(s), b.x(s))
        assert np.abs(a.x(s) - c.x(s)).max() <= 1e-4 * sc.grid.dx


def test_graph_replay_matches_eager():
    sc = _scenes(4, seed=5)
    g = _batch(sc)
    e = _batch(sc)
    e._sim.use_graph = False
    _run(g, ticks=3)
    _run(e, ticks=3)
    for s in range(4):
        assert np.abs(g.x(s) - e.x(s)).max() <= 1e-4 * sc.grid.dx


def test_autotune_restores_state_and_picks_a_candidate():
    sc = _scenes(8, seed=6)
    a = _batch(sc)
    _run(a, ticks=1)
    res = a.autotune(DT, 25, ticks=1, sort_intervals=(1, 2))
    assert set(res["chosen"]) == {"threads_per_workgroup", "particles_per_thread", "paired", "sort_interval"}
    assert a.launch in LAUNCH_SHAPES and a.sort_interval in (1, 2)
    assert "paired512" in res["shapes"]  # the deterministic batch times the paired pass too
    ref = _run(_batch(sc), ticks=2)
    _run(a, ticks=1)
    T = 2 * 25 * DT
    for s in range(8):
        assert np.abs(a.x(s) - ref.x(s)).max() <= 1e-4 * sc.grid.dx
        f_r = ref.tool_force(s, sc.tool[s], T)
        assert np.allclose(a.tool_force(s, sc.tool[s], T), f_r, rtol=1e-3,
                           atol=1e-3 * np.abs(f_r).max() + 1e-9)


@pytest.mark.parametrize("deterministic", [True, False])
def test_scene_mass_conserved_and_isolated(deterministic):
    """Total grid mass of each scene after a scatter equals that scene's particle mass
    (fp32 round-off): nothing leaks between scenes (fixed-point and float scatters)."""
    sc = _scenes(5, seed=7)
    b = _run(_batch(sc, deterministic=deterministic), ticks=1)
    sim = b._sim
    from mpm_amd.core import native
    sim.node_in.zero_()
    sim._pass(native.PASS_STRESS | native.PASS_P2G, DT, sim.stream())
    if deterministic:  # integer accumulators -> the float layout of node_in, for reading
        q = sim.node_fix.view(-1, 4).double()
        sim.node_in.view(-1, 4)[:, 3] = (q[:, 3] / 2.0 ** sim._fix_exp[0]).float()
        sim.node_fix.zero_()
    m, _, _ = b.grid_dense()
    per_scene = m.sum(dim=(1, 2, 3)).cpu().numpy().astype(np.float64)
    mass = sim.mass.cpu().numpy().astype(np.float64)
    scene = sim.scene.cpu().numpy()
    want = np.array([mass[scene == s].sum() for s in range(5)])
    assert np.allclose(per_scene, want, rtol=1e-5)
    sim.node_in.zero_()


def test_bad_scene_id_is_reported():
    """A scene id outside the batch is skipped by the kernels and raised by the guard."""
    sc = _scenes(2, seed=8)
    b = _batch(sc)
    b._sim.scene[0] = 7
    b.step(DT, 5)
    with pytest.raises(RuntimeError, match="scene id"):
        b.step(DT, 5)


def test_single_scene_batch_equals_solver():
    """A batch of one scene is the plain Solver (no sort, so both keep load order)."""
    from mpm_amd import Solver

    sc = _scenes(1, seed=9)
    b = _run(_batch(sc, sort_interval=0))
    s = sc.build_one(Solver(grid=sc.grid, device="hip", sparse=True), 0)
    _run(s)
    assert np.abs(b.x(0) - s.x()).max() <= 1e-4 * sc.grid.dx
    T = 2 * 25 * DT
    f_s = s.tool_force(sc.tool_one, T)
    assert np.allclose(b.tool_force(0, sc.tool[0], T), f_s, rtol=1e-3, atol=1e-3 * np.abs(f_s).max() + 1e-9)
