"""The deterministic scatter mode (Solver(deterministic=True)) on the GPU.

Every scatter contribution is rounded to 64-bit fixed point and accumulated with integer
atomics, and wrench sums the same way, so the grid sums cannot depend on the order the
atomics land in. The checks: two independent runs are bit-identical (positions,
velocities, F, stress, every wrench); the result is bit-identical across launch shapes
(particles per thread) and, as a multiset of particle states, across particle orders
(block sort on / off); it matches the oracle within the parity bar (C1 positions after
100 substeps within 1e-4 dx, colliders and wrenches within 1e-3); and a contribution
beyond the fixed-point range is reported, not wrapped."""
from __future__ import annotations

import numpy as np
import pytest

from mpm_amd import Solver
from mpm_amd.core import native
from oracle.sim import OracleSolver

from test_hip_parity import _c1, _collider_run

pytestmark = pytest.mark.gpu


def _state(s):
    return [s.x(), s.v(), s.F(), s.stress()]


def _collider_det(**kw):
    s, out = _collider_run(Solver, device="hip", deterministic=True, **kw)
    flat = np.concatenate([np.concatenate([c["force"], c["torque"], d["force"], d["torque"], b])
                           for c, d, b in out])
    return s, flat


def test_two_runs_are_bit_identical():
    a, wa = _collider_det()
    b, wb = _collider_det()
    for u, v in zip(_state(a), _state(b)):
        assert np.array_equal(u, v)
    assert np.array_equal(wa, wb)
    assert np.abs(wa).max() > 0


def test_launch_shape_does_not_change_a_bit():
    """4 particles per thread (other chunk-to-workgroup grouping, other atomic order)
    against the default shape."""
    c = _c1(Solver, device="hip", deterministic=True)
    d = _c1(Solver, device="hip", deterministic=True)
    d._sim.launch_bits = native.PASS_PPT4
    for _ in range(5):
        c.step(1e-4, 20)
        d.step(1e-4, 20)
    for u, v in zip(_state(c), _state(d)):
        assert np.array_equal(u, v)


@pytest.mark.parametrize("scene", ["c1", "pour64"])
def test_paired_pass_does_not_change_a_bit(scene):
    """The paired interior pass (two particles per lane over group-aligned chunks of up to
    512, one tile, one arena and one flush for both) against the single-particle pass, on
    cell-sorted particles: every state array and the wrench bitwise equal after 100 (C1,
    the solid specialisation) or 432 (the 64^3 pour, the viscous one) substeps."""
    from mpm_amd.scenes import PourScene

    runs = []
    for pair in (True, False):
        if scene == "c1":
            s = _c1(Solver, device="hip", deterministic=True, sort_interval=1, sort_key="cell")
            s._sim.pair_pass = pair
            for _ in range(5):
                s.step(1e-4, 20)
            w = np.zeros(1)
        else:
            sc = PourScene(n_grid=64)
            s = sc.build(Solver(grid=sc.grid, device="hip", sort_interval=1, sort_key="cell"))
            s._sim.pair_pass = pair
            s.reset_cup_wrench(sc.src)
            sc.frame(s)
            cw = s.cup_wrench(sc.src, sc.dt * sc.substeps)
            w = np.concatenate([cw["force"], cw["torque"]])
        assert s._sim.chunk_start2 is not None and s._sim.chunk_capacity2 > 0
        runs.append((_state(s), w, s._sim.chunk_count.tolist()))
    (a, wa, (n256, n512)), (b, wb, _) = runs
    assert n512 < n256  # fewer, larger chunks
    for u, v in zip(a, b):
        assert np.array_equal(u, v)
    assert np.array_equal(wa, wb)


def test_particle_order_does_not_change_a_bit():
    """Block sort every tick vs never: the same multiset of particle states, bitwise."""
    a = _c1(Solver, device="hip", deterministic=True, sort_interval=1, sort_key="cell")
    b = _c1(Solver, device="hip", deterministic=True, sort_interval=0)
    for _ in range(5):
        a.step(1e-4, 20)
        b.step(1e-4, 20)

    def rows(s):
        r = np.concatenate([s.x(), s.v(), s.F().reshape(-1, 9)], axis=1)
        return r[np.lexsort(r.T[::-1])]

    assert np.array_equal(rows(a), rows(b))


def test_matches_oracle_within_parity_bar():
    hip = _c1(Solver, device="hip", deterministic=True)
    ora = _c1(OracleSolver)
    for _ in range(5):
        hip.step(1e-4, 20)
        ora.step(1e-4, 20)
    dx = 1.0 / 32
    assert np.abs(hip.x() - ora.x()).max() <= 1e-4 * dx
    s, wh = _collider_run(Solver, device="hip", deterministic=True)
    o, wo = _collider_run(OracleSolver)
    assert np.abs(s.x() - o.x()).max() <= 1e-4 * 0.01
    for (ch, sh, bh), (co, so, bo) in zip(wh, wo):
        for a, b in ((ch["force"], co["force"]), (ch["torque"], co["torque"]),
                     (sh["force"], so["force"]), (sh["torque"], so["torque"]), (bh, bo)):
            np.testing.assert_allclose(a, b, rtol=1e-3, atol=1e-4 * np.abs(b).max())


def test_sparse_deterministic_is_bit_identical_and_equals_dense():
    """Sparse block lists come from an atomic compaction (order varies run to run); the
    fixed-point sums make the result independent of it, and of dense vs sparse."""
    a = _c1(Solver, device="hip", deterministic=True, sparse=True)
    b = _c1(Solver, device="hip", deterministic=True, sparse=True)
    c = _c1(Solver, device="hip", deterministic=True)
    for _ in range(5):
        for s in (a, b, c):
            s.step(1e-4, 20)
    for u, v, w in zip(_state(a), _state(b), _state(c)):
        assert np.array_equal(u, v) and np.array_equal(u, w)


def test_range_overflow_is_reported():
    """A scale far too fine for the state (forced from the host) drops contributions
    beyond 2^51 units and raises at the next guard instead of wrapping silently."""
    s = _c1(Solver, device="hip", deterministic=True)
    s.step(1e-4, 2)
    sim = s._sim

    def too_fine(dt):
        sim._fix_exp = (120, 120)
        sim._structs()

    sim._fix_scales = too_fine
    s.step(1e-4, 2)
    with pytest.raises(RuntimeError, match="fixed-point range"):
        s.step(1e-4, 2)


def test_cdf_is_rejected():
    from mpm_amd.geometry import build_surface_cdf, square_sheet

    s = _c1(Solver, device="hip", deterministic=True)
    v, f = square_sheet(0.3)
    with pytest.raises(NotImplementedError, match="deterministic"):
        s.add_cdf_collider(build_surface_cdf(v, f, res=24), center=(0.5, 0.5, 0.3))
