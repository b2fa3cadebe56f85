"""CPIC thin-boundary (CDF) colliders on the HIP path against the oracle (GPU).

The HIP stamps must reproduce the oracle's node tags and owner distances; trajectories
with a shelf, a dam and a moving wall must follow the oracle's within 1e-4 dx over the
checked horizon; the physical claims of the reference's CDF tests (nothing crosses the
sheet, nothing is left behind the moving wall, the shelf reaction reads a fraction of
the weight) must hold on the GPU too; fused and split pipelines, mask on / off and graph
replay versus eager must agree. Float atomics (grid scatter and reaction sums) and FMA
contraction in the device trilinear interpolation are the only expected sources of
difference."""
from __future__ import annotations

import numpy as np
import pytest

from mpm_amd import GridConfig, Solver
from mpm_amd.geometry import build_surface_cdf, square_sheet
from mpm_amd.materials import elastic, newtonian
from oracle.sim import OracleSolver

pytestmark = pytest.mark.gpu

G = GridConfig(n_grid=32, grid_lim=0.4)
DX = G.dx
DT, SUB = 2e-4, 10
_SHEETS = {}


def sheet(normal="z", size=0.5, band_cells=3.0):
    key = (normal, size, band_cells)
    if key not in _SHEETS:
        v, f = square_sheet(size, normal_axis={"x": 0, "y": 1, "z": 2}[normal])
        _SHEETS[key] = build_surface_cdf(v, f, res=64, band_cells=band_cells)
    return _SHEETS[key]


def water(x0, x1, y0, y1, z0, z1):
    hp = DX / 2
    ax = [np.arange(a + hp / 2, b, hp) for a, b in ((x0, x1), (y0, y1), (z0, z1))]
    return np.stack(np.meshgrid(*ax, indexing="ij"), -1).reshape(-1, 3).astype(np.float32)


def make(cls, pos, g=(0.0, 0.0, 0.0), fused=True, plane=True, material=None, **kw):
    s = cls(grid=G, fused=fused, **kw) if cls is Solver else cls(G, fused=fused, **kw)
    s.load_particles(pos, np.full(len(pos), (DX / 2) ** 3, np.float32))
    s.set_material(material or newtonian(eta=0.5, density=1000.0, bulk_modulus=5.0e4))
    if cls is Solver:
        s._sim.set_gravity(g)
    else:
        s._sim.gravity[:] = g
    if plane:
        s.add_plane((0, 0, 3 * DX), (0, 0, 1), "slip")
    s.add_domain_walls()
    return s


def hip(pos, **kw):
    return make(Solver, pos, device="hip", **kw)


def orc(pos, **kw):
    return make(OracleSolver, pos, **kw)


def dense(sim, t):
    """Blocked node array -> dense (n, n, n), z fastest."""
    n = sim.n_grid
    i, j, k = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    nb = sim.nb
    idx = ((((i >> 2) * nb[1] + (j >> 2)) * nb[2] + (k >> 2)) << 6) + ((i & 3) << 4) + ((j & 3) << 2) + (k & 3)
    return t.cpu().numpy()[idx]


def shelf(cls, fused=True, moving=False, two_blobs=False):
    pos = water(0.14, 0.26, 0.14, 0.26, 0.21, 0.28)
    if two_blobs:
        pos = np.concatenate([pos, water(0.06, 0.12, 0.06, 0.12, 0.08, 0.14)])
    s = (hip if cls is Solver else orc)(pos, g=(0.0, 0.0, -9.81), fused=fused)
    h = s.add_cdf_collider(sheet("z"), center=(0.2, 0.2, 0.20), friction=0.3,
                           velocity=(0.0, 0.0, 0.01) if moving else (0.0, 0.0, 0.0))
    return s, h


def test_stamp_tags_and_distances_match_oracle():
    """Two lanes (a tilted sheet and a vertical one) stamped by the HIP kernels against
    the oracle's stamps: tags identical up to a handful of nodes whose interpolated
    distance sits within fp32 rounding (FMA contraction on the device) of a band edge or
    of zero, owner distances within 1e-6 m, occupancy mask identical."""
    pos = water(0.08, 0.12, 0.08, 0.12, 0.08, 0.12)
    a, b = hip(pos), orc(pos)
    q = (0.0, np.sin(0.2), 0.0, np.cos(0.2))
    for s in (a, b):
        s.add_cdf_collider(sheet("z", band_cells=6.0), center=(0.2, 0.2, 0.2), quat=q, band=2.5 * DX)
        s.add_cdf_collider(sheet("x"), center=(0.26, 0.2, 0.2))
    a._sim.cdf.stamp(0.0, lag_prev=False)
    b._sim._stamp_cdf(DT)
    ta, tb = dense(a._sim, a._sim.cdf.tag), b._sim.cdf_tag
    da, db = dense(a._sim, a._sim.cdf.dist), b._sim.cdf_d
    assert (tb != 0).sum() > 1000
    diff = ta != tb
    assert int(diff.sum()) <= max(3, int(0.002 * (tb != 0).sum())), f"{int(diff.sum())} tags differ"
    both = (ta != 0) & (tb != 0) & ~diff
    assert np.allclose(da[both], db[both], atol=1e-6)
    assert np.array_equal(dense(a._sim, a._sim.cdf.tag_prev), ta)
    blk = a._sim.cdf.block.cpu().numpy().reshape(8, 8, 8)
    occ = ta.reshape(8, 4, 8, 4, 8, 4).transpose(0, 2, 4, 1, 3, 5).reshape(8, 8, 8, 64)
    assert np.array_equal(blk, np.bitwise_or.reduce(occ, axis=-1))
    assert np.array_equal(blk != 0, b._sim.cdf_block != 0)


@pytest.mark.parametrize("fused", [True, False])
def test_shelf_trajectory_matches_oracle(fused):
    """Column falling onto a horizontal sheet: 5 ticks of 10 substeps, positions within
    1e-4 dx of the oracle and the particle tags identical."""
    a, _ = shelf(Solver, fused=fused)
    b, _ = shelf(OracleSolver, fused=fused)
    for _ in range(5):
        a.step(DT, SUB)
        b.step(DT, SUB)
    err = np.abs(a.x() - b.x()).max()
    assert err <= 1e-4 * DX, f"max |dx| {err / DX:.2e} dx"
    assert np.array_equal(a._sim.cdf_tag.cpu().numpy(), b._sim.particle_cdf_tag)


def test_moving_sheet_trajectory_matches_oracle():
    """A sheet rising at 1 cm/s under the column (the prev / cur double buffers)."""
    a, _ = shelf(Solver, moving=True)
    b, _ = shelf(OracleSolver, moving=True)
    for _ in range(4):
        a.step(DT, SUB)
        b.step(DT, SUB)
    assert np.abs(a.x() - b.x()).max() <= 1e-4 * DX
    assert a._sim.cdf.stamp_count == b._sim._cdf_stamp_count == 4 * SUB


def test_dam_holds_and_matches_oracle():
    """A slab driven at 1 m/s into a zero-thickness vertical sheet: no particle crosses
    (tests/test_cdf_transfers.py:43-62), and the first ticks follow the oracle."""
    pos = water(0.10, 0.16, 0.12, 0.28, 3 * DX, 3 * DX + 0.08)
    a, b = hip(pos), orc(pos)
    for s in (a, b):
        s.add_cdf_collider(sheet("x"), center=(0.20, 0.2, 0.2))
    v = a.v()
    v[:, 0] += 1.0
    a.set_v(v)
    b._sim.v[:, 0] += 1.0
    for _ in range(3):
        a.step(DT, SUB)
        b.step(DT, SUB)
    assert np.abs(a.x() - b.x()).max() <= 1e-4 * DX
    crossed, near = 0, False
    for _ in range(27):
        a.step(DT, SUB)
        x = a.x()
        crossed = max(crossed, int((x[:, 0] > 0.20).sum()))
        near = near or bool((x[:, 0] > 0.20 - 1.5 * DX).any())
    assert near and crossed == 0, (near, crossed)


def test_moving_wall_leaves_nothing_behind():
    """tests/test_cdf_transfers.py:155-173 on the GPU."""
    s = hip(water(0.22, 0.30, 0.14, 0.26, 3 * DX, 3 * DX + 0.08))
    s.add_cdf_collider(sheet("x"), center=(0.18, 0.2, 0.2), velocity=(1.0, 0.0, 0.0))
    t = 0.0
    for _ in range(40):
        s.step(DT, SUB)
        t += DT * SUB
        assert int((s.x()[:, 0] < 0.18 + t - 1.5 * DX).sum()) == 0


def test_shelf_wrench_matches_oracle():
    """Column landing on the shelf: the blocked-deposit reaction of every tick matches
    the oracle's within 1e-3 of the column weight."""
    a, ha = shelf(Solver)
    b, hb = shelf(OracleSolver)
    weight = float(a.n_particles * (DX / 2) ** 3 * 1000.0 * 9.81)
    for _ in range(25):  # 1 cm of free fall onto the sheet
        a.step(DT, SUB)
        b.step(DT, SUB)
    fa, fb = [], []
    for _ in range(10):
        a.reset_cdf_wrench()
        b.reset_cdf_wrench()
        a.step(DT, SUB)
        b.step(DT, SUB)
        fa.append(a.cdf_wrench(ha, DT * SUB)["force"])
        fb.append(b.cdf_wrench(hb, DT * SUB)["force"])
    fa, fb = np.array(fa), np.array(fb)
    assert np.abs(fb).max() > 0.1 * weight, "the column never loaded the shelf"
    assert np.abs(fa - fb).max() <= 1e-3 * weight, (fa, fb)


def test_settled_shelf_wrench_reads_a_fraction_of_the_weight():
    """tests/test_cdf_transfers.py:65-100 on the GPU: after settling, the mean reaction
    points down and reads 0.15-0.6 of the weight; nothing fell through."""
    a, ha = shelf(Solver)
    weight = float(a.n_particles * (DX / 2) ** 3 * 1000.0 * 9.81)
    for _ in range(120):
        a.step(DT, SUB)
    a._sim.model[0]["grid_v_damping_scale"] = 0.999
    a._sim._model_dirty = True
    forces = []
    for _ in range(40):
        a.reset_cdf_wrench()
        a.step(DT, SUB)
        forces.append(a.cdf_wrench(ha, DT * SUB)["force"][2])
    f = float(np.mean(forces))
    assert f < 0.0 and 0.15 * weight < -f < 0.6 * weight, (f, weight)
    assert float(a.x()[:, 2].min()) > 0.20 - 1.5 * DX


def test_fused_split_mask_and_graph_agree():
    """Static shelf: fused vs split, block pretest on vs off, graph replay vs eager."""
    ref, _ = shelf(Solver, two_blobs=True)
    runs = {"ref": ref}
    runs["split"], _ = shelf(Solver, fused=False, two_blobs=True)
    runs["mask_off"], _ = shelf(Solver, two_blobs=True)
    runs["mask_off"]._sim.cdf.mask_off = 1
    runs["eager"], _ = shelf(Solver, two_blobs=True)
    runs["eager"]._sim.use_graph = False
    for _ in range(6):
        for s in runs.values():
            s.step(DT, SUB)
    for name, s in runs.items():
        assert np.abs(s.x() - ref.x()).max() <= 1e-4 * DX, name
    assert ref._sim._graph is not None, "static CDF ticks should replay a captured graph"
    assert ref._sim.cdf.stamp_count == 1


def test_friction_slows_sliding_block():
    """tests/test_cdf_transfers.py:121-152 on the GPU: frictionless slides, mu = 2 drags."""
    pos = water(0.16, 0.24, 0.16, 0.24, 0.21, 0.26)
    slid = {}
    for mu in (0.0, 2.0):
        s = hip(pos, g=(4.0, 0.0, -9.0), plane=False, material=elastic(E=5.0e4, nu=0.3, density=1000.0))
        s.add_cdf_collider(sheet("z"), center=(0.2, 0.2, 0.20), surface="separable", friction=mu)
        for _ in range(30):
            s.step(DT, SUB)
        vels = []
        for _ in range(10):
            s.step(DT, SUB)
            vels.append(float(s.v()[:, 0].mean()))
        slid[mu] = float(np.mean(vels))
    assert slid[0.0] > 1.5 * max(slid[2.0], 1e-6) and slid[0.0] > 0.01, slid


def test_cdf_rejects_lean_pass_and_batches():
    from mpm_amd.core import native
    s, _ = shelf(Solver)
    s.step(DT, 2)
    sim = s._sim
    with pytest.raises(native.NativeError):
        sim._pass(native.PASS_G2P | native.PASS_STRESS | native.PASS_P2G | native.PASS_LEAN, DT, sim.stream())
