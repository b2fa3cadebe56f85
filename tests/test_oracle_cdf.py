"""CPIC thin-boundary (CDF) collider in the oracle and the geometry builder (CPU).

Restates the reference's CDF tests against the oracle: stamp bits against an analytic
plane (tests/test_cdf_collider.py:40-73), the guards (:76-88), an inert collider being
bitwise invisible (:91-104), the zero-thickness dam (tests/test_cdf_transfers.py:43-62),
the shelf reaction reading a scaled fraction of the weight (:65-100), friction through the
ghost projection (:121-152), a moving wall that leaves nothing behind (:155-173), the
sweep warning (:176-187), fused == split bitwise static and moving
(tests/test_cdf_fused.py:37-47), the block-mask skip being exact
(tests/test_cdf_mask.py:49-75) and the static-pose stamp skip being exact
(tests/test_cdf_stamp_skip.py:68-92). The reference's own expectations are the pins here:
physical outcomes (no particle crosses, weight fraction window, monotone friction) and
bitwise equalities between paths that must agree. The geometry builder is pinned against
closed-form distances of a flat sheet and a cylinder (tests/test_cdf_geometry.py)."""
from __future__ import annotations

import warnings

import numpy as np
import pytest

from mpm_amd import GridConfig
from mpm_amd.geometry import (CDFData, build_surface_cdf, load_cdf, open_edges,
                              revolve_profile_open, save_cdf, square_sheet)
from mpm_amd.materials import elastic, newtonian
from oracle.sim import MAX_CDF, OracleSolver

G = GridConfig(n_grid=32, grid_lim=0.4)
DX = G.dx
DT, SUB = 2e-4, 10


def _sheet(normal="z", size=0.5, res=64, band_cells=3.0):
    v, f = square_sheet(size, normal_axis={"x": 0, "y": 1, "z": 2}[normal])
    return build_surface_cdf(v, f, res=res, band_cells=band_cells)


_CACHE = {}


def sheet(normal="z", size=0.5, res=64, band_cells=3.0) -> CDFData:
    key = (normal, size, res, band_cells)
    if key not in _CACHE:
        _CACHE[key] = _sheet(normal, size, res, band_cells)
    return _CACHE[key]


def water(x0, x1, y0, y1, z0, z1):
    hp = DX / 2
    ax = [np.arange(a + hp / 2, b, hp) for a, b in ((x0, x1), (y0, y1), (z0, z1))]
    return np.stack(np.meshgrid(*ax, indexing="ij"), -1).reshape(-1, 3).astype(np.float32)


def solver(pos, g=(0.0, 0.0, 0.0), fused=True, plane=True, material=None):
    s = OracleSolver(G, fused=fused)
    s.load_particles(pos, np.full(len(pos), (DX / 2) ** 3, np.float32))
    s.set_material(material or newtonian(eta=0.5, density=1000.0, bulk_modulus=5.0e4))
    s._sim.gravity[:] = g
    if plane:
        s.add_plane((0, 0, 3 * DX), (0, 0, 1), "slip")
    s.add_domain_walls()
    return s


# --------------------------------------------------------------------------- geometry
def test_sheet_cdf_signed_distance_and_rim():
    """Flat sheet: |d| equals the plane distance, sign follows the normal, the rim tube
    and the far field are invalid (tests/test_cdf_geometry.py)."""
    cdf = sheet("z", size=0.3, band_cells=6.0)
    h = cdf.cell
    ax = cdf.origin[0] + h * np.arange(cdf.res)
    z = cdf.origin[2] + h * np.arange(cdf.res)
    inner = np.abs(ax) < 0.1
    for dz in (-3, -1, 1, 3):
        k = int(np.argmin(np.abs(z - dz * h)))
        d = cdf.values[np.ix_(inner, inner, [k])]
        ok = cdf.valid[np.ix_(inner, inner, [k])]
        assert np.all(ok == 1.0)
        assert np.allclose(d, z[k], atol=0.15 * h)
    rim = int(np.argmin(np.abs(ax - 0.15)))
    assert np.all(cdf.valid[rim, inner, :] == 0.0)
    assert np.all(cdf.valid[:, :, 0] == 0.0)


def test_cylinder_cdf_side_and_open_edges():
    v, f = revolve_profile_open([(0.1, 0.0), (0.1, 0.2)], n_theta=48)
    assert len(open_edges(f)) == 2 * 48
    cdf = build_surface_cdf(v, f, res=32, band_cells=3.0)
    h = cdf.cell
    zc = int(round((0.1 - cdf.origin[2]) / h))
    xs = cdf.origin[0] + h * np.arange(cdf.res)
    yc = int(round((0.0 - cdf.origin[1]) / h))
    r = np.abs(xs)
    band = (np.abs(r - 0.1) < 2 * h) & (np.abs(r - 0.1) > 0.3 * h)
    d = cdf.values[band, yc, zc]
    assert np.all(cdf.valid[band, yc, zc] == 1.0)
    assert np.all(np.sign(d) == np.sign(r[band] - 0.1))
    assert np.allclose(np.abs(d), np.abs(r[band] - 0.1), atol=0.2 * h)


def test_cdf_save_load_roundtrip(tmp_path):
    cdf = sheet("x", size=0.2, res=24)
    p = tmp_path / "sheet.npz"
    save_cdf(cdf, p)
    back = load_cdf(p)
    assert np.array_equal(back.values, cdf.values) and np.array_equal(back.valid, cdf.valid)
    assert back.cell == cdf.cell and back.band == cdf.band and back.res == 24


# ----------------------------------------------------------------------------- stamps
def test_stamp_bits_match_analytic_plane():
    """tests/test_cdf_collider.py:40-73: valid and side bits above / below a horizontal
    sheet, nothing beyond the band or in the rim tube, owner distance = |z - z0|."""
    s = OracleSolver(G, fused=False)
    rng = np.random.default_rng(0)
    s.load_particles(np.full((8, 3), 0.2, np.float32) + rng.uniform(0, 0.01, (8, 3)).astype(np.float32),
                     np.full(8, (DX / 2) ** 3, np.float32))
    s.set_material(newtonian(eta=0.2, density=1000.0, bulk_modulus=5.0e4))
    z0 = 0.2
    h = s.add_cdf_collider(sheet("z", size=0.3, band_cells=6.0), center=(0.2, 0.2, z0), band=2.5 * DX)
    assert h == 0
    s._sim._stamp_cdf(1e-4)
    tag = s._sim.cdf_tag
    k0 = z0 / DX
    ij = np.arange(int(0.15 / DX), int(0.25 / DX))
    above = tag[np.ix_(ij, ij, [int(k0 + 1), int(k0 + 2)])]
    below = tag[np.ix_(ij, ij, [int(k0 - 1), int(k0 - 2)])]
    assert np.all(above & 1) and np.all(above & 2)
    assert np.all(below & 1) and np.all(below & 2 == 0)
    assert np.all(tag[np.ix_(ij, ij, [int(k0 + 5), int(k0 - 5)])] == 0)
    assert np.all(tag[int(0.05 / DX), ij, int(k0 + 1)] == 0)
    d = s._sim.cdf_d
    assert np.allclose(d[np.ix_(ij, ij, [int(k0 + 1)])], abs(int(k0 + 1) * DX - z0), atol=0.1 * DX)
    # the block mask marks exactly the 4^3 blocks holding a tag (cur | prev)
    occ = (tag | s._sim.cdf_tag_prev).reshape(8, 4, 8, 4, 8, 4).transpose(0, 2, 4, 1, 3, 5)
    want = np.bitwise_or.reduce(occ.reshape(8, 8, 8, 64), axis=-1)
    assert np.array_equal(s._sim.cdf_block, want)


def test_owner_is_the_nearest_lane():
    """Two parallel sheets 3 dx apart: nodes between them carry both lanes' valid bits and
    the owner bits name the nearer sheet."""
    s = solver(water(0.05, 0.07, 0.05, 0.07, 0.05, 0.07))
    cdf = sheet("z", band_cells=6.0)
    a = s.add_cdf_collider(cdf, center=(0.2, 0.2, 0.2), band=2.6 * DX)
    b = s.add_cdf_collider(cdf, center=(0.2, 0.2, 0.2 + 3 * DX), band=2.6 * DX)
    s._sim._stamp_cdf(DT)
    tag = s._sim.cdf_tag
    k0 = int(round(0.2 / DX))
    t1, t2 = tag[16, 16, k0 + 1], tag[16, 16, k0 + 2]
    for t in (t1, t2):
        assert (t >> (2 * a)) & 1 and (t >> (2 * b)) & 1
        assert (t >> (2 * a + 1)) & 1 == 1 and (t >> (2 * b + 1)) & 1 == 0
    assert (t1 >> 8) & 3 == a and (t2 >> 8) & 3 == b


def test_guards():
    """tests/test_cdf_collider.py:76-88."""
    s = solver(water(0.08, 0.14, 0.08, 0.14, 0.08, 0.14))
    cdf = sheet("z", size=0.3, band_cells=6.0)
    with pytest.raises(ValueError, match="runtime band"):
        s.add_cdf_collider(cdf, center=(0.2, 0.2, 0.2), band=10.0)
    with pytest.warns(RuntimeWarning, match="below 1.5 dx"):
        h = s.add_cdf_collider(cdf, center=(0.2, 0.2, 0.2), band=0.5 * DX)
    with pytest.warns(RuntimeWarning, match="jumped the surface"):
        s.set_cdf_pose(h, center=(0.2, 0.2, 0.35))
    for _ in range(MAX_CDF - 1):
        s.add_cdf_collider(cdf, center=(0.2, 0.2, 0.2))
    with pytest.raises(ValueError, match="at most"):
        s.add_cdf_collider(cdf, center=(0.2, 0.2, 0.2))


@pytest.mark.parametrize("fused", [False, True])
def test_inert_collider_is_bitwise_invisible(fused):
    """tests/test_cdf_collider.py:91-104: a sheet far from the material changes nothing."""
    pos = water(0.08, 0.14, 0.08, 0.14, 0.08, 0.14)
    a = solver(pos, fused=fused)
    b = solver(pos, fused=fused)
    b.add_cdf_collider(sheet("z", size=0.3, band_cells=6.0), center=(0.30, 0.30, 0.30))
    for _ in range(4):
        a.step(DT, SUB)
        b.step(DT, SUB)
    assert np.array_equal(a.x(), b.x()) and np.array_equal(a.v(), b.v()) and np.array_equal(a.F(), b.F())


# -------------------------------------------------------------------------- transfers
def test_zero_thickness_dam_holds():
    """tests/test_cdf_transfers.py:43-62: a slab driven at 1 m/s into a zero-thickness
    vertical sheet never crosses it."""
    s = solver(water(0.10, 0.16, 0.12, 0.28, 3 * DX, 3 * DX + 0.08))
    x_wall = 0.20
    s.add_cdf_collider(sheet("x"), center=(x_wall, 0.2, 0.2))
    s._sim.v[:, 0] += 1.0
    crossed, near = 0, False
    for _ in range(30):
        s.step(DT, SUB)
        x = s.x()
        crossed = max(crossed, int((x[:, 0] > x_wall).sum()))
        near = near or bool((x[:, 0] > x_wall - 1.5 * DX).any())
    assert near, "the slab never reached the wall"
    assert crossed == 0, f"{crossed} particles crossed the thin wall"


def test_shelf_wrench_reads_a_scaled_fraction_of_the_weight():
    """tests/test_cdf_transfers.py:65-100: a column settled on a horizontal sheet; the
    blocked-deposit reaction points down and reads 0.15-0.6 of the weight."""
    pos = water(0.14, 0.26, 0.14, 0.26, 0.21, 0.29)
    s = solver(pos, g=(0.0, 0.0, -9.81))
    h = s.add_cdf_collider(sheet("z"), center=(0.2, 0.2, 0.20))
    weight = float(len(pos) * (DX / 2) ** 3 * 1000.0 * 9.81)
    for _ in range(120):
        s.step(DT, SUB)
    s._sim.grid_v_damping_scale = 0.999
    forces = []
    for _ in range(40):
        s.reset_cdf_wrench()
        s.step(DT, SUB)
        forces.append(s.cdf_wrench(h, DT * SUB)["force"][2])
    f_mean = float(np.mean(forces))
    assert f_mean < 0.0
    assert 0.15 * weight < -f_mean < 0.6 * weight, (f_mean, weight)
    assert float(s.x()[:, 2].min()) > 0.20 - 1.5 * DX


def test_ghost_friction_slides_and_holds():
    """tests/test_cdf_transfers.py:121-152: an elastic block on a shelf under tilted
    gravity slides frictionless and is slowed by friction 2.0."""
    pos = water(0.16, 0.24, 0.16, 0.24, 0.21, 0.26)
    slid = {}
    for mu, key in ((0.0, "slip"), (2.0, "hold")):
        s = solver(pos, g=(4.0, 0.0, -9.0), plane=False, material=elastic(E=5.0e4, nu=0.3, density=1000.0))
        s.add_cdf_collider(sheet("z"), center=(0.2, 0.2, 0.20), surface="separable", friction=mu)
        for _ in range(30):
            s.step(DT, SUB)
        vels = []
        for _ in range(10):
            s.step(DT, SUB)
            vels.append(float(s.v()[:, 0].mean()))
        slid[key] = float(np.mean(vels))
    assert slid["slip"] > 1.5 * max(slid["hold"], 1e-6), slid
    assert slid["slip"] > 0.01


def test_moving_wall_sweeps_without_leaks():
    """tests/test_cdf_transfers.py:155-173: a vertical sheet driven at 1 m/s through a
    resting blob leaves no particle behind it."""
    s = solver(water(0.22, 0.30, 0.14, 0.26, 3 * DX, 3 * DX + 0.08))
    x0, v_wall, t = 0.18, 1.0, 0.0
    s.add_cdf_collider(sheet("x"), center=(x0, 0.2, 0.2), velocity=(v_wall, 0.0, 0.0))
    for _ in range(40):
        s.step(DT, SUB)
        t += DT * SUB
        behind = int((s.x()[:, 0] < x0 + v_wall * t - 1.5 * DX).sum())
        assert behind == 0, f"{behind} particles tunneled behind the moving wall"
    assert x0 + v_wall * t > 0.25


def test_fast_wall_warns_on_band_sweep():
    """tests/test_cdf_transfers.py:176-187."""
    s = solver(water(0.14, 0.20, 0.14, 0.26, 3 * DX, 3 * DX + 0.05))
    s.add_cdf_collider(sheet("x"), center=(0.3, 0.2, 0.2), velocity=(2.0 * 2.0 * DX / DT, 0.0, 0.0))
    with pytest.warns(RuntimeWarning, match="sweeps"):
        s.step(DT, 1)


# ------------------------------------------------------------------ path equivalences
def _shelf_scene(fused, moving=False, two_blobs=False, mask_off=False, force=False):
    pos = water(0.14, 0.26, 0.14, 0.26, 0.21, 0.28)
    if two_blobs:
        pos = np.concatenate([pos, water(0.06, 0.12, 0.06, 0.12, 0.08, 0.14)])
    s = solver(pos, g=(0.0, 0.0, -9.81), fused=fused)
    s.add_cdf_collider(sheet("z"), center=(0.2, 0.2, 0.20), friction=0.3,
                       velocity=(0.0, 0.0, 0.01) if moving else (0.0, 0.0, 0.0))
    s._sim.cdf.mask_off = 1 if mask_off else 0
    s._sim._cdf_force_stamp = force
    return s


@pytest.mark.parametrize("moving", [False, True])
def test_fused_matches_split_bitwise(moving):
    """tests/test_cdf_fused.py:37-47: the prev / cur tag and lane double buffers make the
    fused pass read exactly what the split pipeline reads."""
    a = _shelf_scene(False, moving)
    b = _shelf_scene(True, moving)
    for _ in range(6):
        a.step(DT, SUB)
        b.step(DT, SUB)
    for fa, fb in ((a.x, b.x), (a.v, b.v), (a.F, b.F), (a.stress, b.stress)):
        assert np.array_equal(fa(), fb())


@pytest.mark.parametrize("fused", [False, True])
def test_block_mask_skip_is_exact(fused):
    """tests/test_cdf_mask.py:49-75: the block pretest changes no bit, static and moving."""
    for moving in (False, True):
        out = {}
        for mask_off in (False, True):
            s = _shelf_scene(fused, moving, two_blobs=True, mask_off=mask_off)
            for _ in range(5):
                s.step(DT, SUB)
            out[mask_off] = (s.x(), s.v())
        assert np.array_equal(out[False][0], out[True][0])
        assert np.array_equal(out[False][1], out[True][1])


@pytest.mark.parametrize("fused", [False, True])
def test_static_stamp_skip_is_exact(fused):
    """tests/test_cdf_stamp_skip.py:68-79: a static shelf stamps once; always restamping
    gives the same bits."""
    out = {}
    for force in (False, True):
        s = _shelf_scene(fused, force=force)
        for _ in range(5):
            s.step(DT, SUB)
        out[force] = (s.x(), s.v(), s._sim._cdf_stamp_count)
    assert np.array_equal(out[False][0], out[True][0]) and np.array_equal(out[False][1], out[True][1])
    assert out[False][2] == 1 and out[True][2] == 5 * SUB


@pytest.mark.parametrize("fused", [False, True])
def test_stop_start_stamp_skip_is_exact(fused):
    """tests/test_cdf_stamp_skip.py:48-65,82-92: sweep, hold, sweep again."""
    out = {}
    for force in (False, True):
        s = solver(water(0.22, 0.30, 0.14, 0.26, 3 * DX, 3 * DX + 0.08), fused=fused)
        h = s.add_cdf_collider(sheet("x"), center=(0.18, 0.2, 0.2), velocity=(1.0, 0.0, 0.0))
        s._sim._cdf_force_stamp = force
        for _ in range(4):
            s.step(DT, SUB)
        s.set_cdf_pose(h, velocity=(0.0, 0.0, 0.0))
        for _ in range(2):
            s.step(DT, SUB)
        s.set_cdf_pose(h, velocity=(0.5, 0.0, 0.0))
        for _ in range(3):
            s.step(DT, SUB)
        out[force] = (s.x(), s.v(), s._sim._cdf_stamp_count)
    assert np.array_equal(out[False][0], out[True][0]) and np.array_equal(out[False][1], out[True][1])
    assert out[False][2] < out[True][2] == 9 * SUB


def test_cdf_planted_mistake_is_caught():
    """The dam test has teeth: with the scatter mask switched off (every lane reads as
    compatible) material crosses the sheet."""
    s = solver(water(0.10, 0.16, 0.12, 0.28, 3 * DX, 3 * DX + 0.08))
    s.add_cdf_collider(sheet("x"), center=(0.20, 0.2, 0.2))
    s._sim.v[:, 0] += 1.0
    s.step(DT, 1)
    s._sim.cdf.n_cdf = 0     # transfers ignore the tags from here on
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        for _ in range(30):
            s.step(DT, SUB)
    assert int((s.x()[:, 0] > 0.20).sum()) > 0
