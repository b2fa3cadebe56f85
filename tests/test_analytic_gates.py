"""The reference's analytic physics gates (tests/test_analytic_benchmarks.py, test_pour.py,
test_sdf_collider.py), run on the oracle (CPU) and on the HIP Solver (GPU) with the
reference's scenes and tolerances:

* sliding block on a frictional incline: a = g (sin th - mu cos th) within 15 %, and the
  mu > tan th block holds (test_analytic_benchmarks.py:60-90): pins the separable plane's
  Coulomb branch end to end;
* hydrostatic column weighed by a box scale inside a slip-walled shaft, settled with grid
  damping 0.999: the box impulse reads rho V g within 8 % (:92-142): pins slip planes, the
  box impulse and the damping branch;
* free-free elastic bar: period 2 L / sqrt(E / rho) within 3 % (:147-191): pins the
  stress, P2G and G2P wave speed;
* water in a revolved-SDF glass: the cup wrench reads the water's weight within 10 % and
  nothing enters the glass solid (test_pour.py:39-57): pins the cup's grab / Coulomb
  branches and its wrench;
* fluid in a mesh-SDF cup: |Fz| of the SDF wrench within 25 % of the weight
  (test_sdf_collider.py:67-80): pins the mesh-SDF response and wrench;
* a block thrown at the domain walls stays inside and keeps sliding along them (the
  reference's bounding box stops outward motion only, mpm_solver_warp.py:2294-2359).

The two heaviest oracle runs (column, cup) use the oracle's OpenMP build: the same
restated kernels with atomic scatter, whose summation order cannot move a physics gate.
The friction-sign mutant of the oracle fails the incline gate (the gate has teeth)."""
from __future__ import annotations

import contextlib

import numpy as np
import pytest

from mpm_amd import GridConfig, Solver
from mpm_amd.materials import elastic, newtonian
from oracle import sim as O
from oracle.sim import OracleSolver

GRAV = 9.81
THETA = np.deg2rad(30.0)
G32 = GridConfig(n_grid=32, grid_lim=0.4)

BACKENDS = [pytest.param("oracle", id="oracle"), pytest.param("hip", id="hip", marks=pytest.mark.gpu)]


def _solver(backend, grid):
    return OracleSolver(grid) if backend == "oracle" else Solver(grid=grid, device="hip")


@contextlib.contextmanager
def _threaded(backend):
    if backend == "oracle":
        O.use_threaded_build(True)
    try:
        yield
    finally:
        if backend == "oracle":
            O.use_threaded_build(False)


def _set_damping(s, scale):
    sim = s._sim
    if isinstance(s, OracleSolver):
        sim.grid_v_damping_scale = scale
    else:
        sim.model[0]["grid_v_damping_scale"] = scale
        sim._model_dirty = True


def _fill(x0, x1, y0, y1, z0, z1, h):
    ax = [np.arange(a + h / 2, b, h) for a, b in ((x0, x1), (y0, y1), (z0, z1))]
    return np.stack(np.meshgrid(*ax, indexing="ij"), -1).reshape(-1, 3).astype(np.float32)


# ---------------------------------------------------------------- 1. incline
def _incline(backend, mu):
    dx = G32.dx
    h = dx / 2
    z0 = 3 * dx
    pos = _fill(0.10, 0.18, 0.16, 0.24, z0, z0 + 0.04, h)
    s = _solver(backend, G32)
    s.load_particles(pos, np.full(len(pos), h ** 3, np.float32))
    s.set_material(elastic(E=1.0e6, nu=0.3, density=1000.0),
                   g=[GRAV * np.sin(THETA), 0.0, -GRAV * np.cos(THETA)])
    s.add_plane((0, 0, z0), (0, 0, 1), "separate", friction=mu)
    return s


def _incline_accel(backend, mu=0.3):
    s = _incline(backend, mu)
    dt, sub = 2e-4, 10
    for _ in range(20):
        s.step(dt, sub)
    v1 = float(s.v()[:, 0].mean())
    n = 50
    for _ in range(n):
        s.step(dt, sub)
    v2 = float(s.v()[:, 0].mean())
    return (v2 - v1) / (n * dt * sub)


@pytest.mark.parametrize("backend", BACKENDS)
def test_incline_sliding_block_matches_coulomb(backend):
    mu = 0.3
    a_true = GRAV * (np.sin(THETA) - mu * np.cos(THETA))
    a = _incline_accel(backend, mu)
    assert abs(a - a_true) / a_true < 0.15, (a, a_true)


def test_incline_gate_has_teeth():
    a_true = GRAV * (np.sin(THETA) - 0.3 * np.cos(THETA))
    with O.mutant(1):  # friction pushes instead of resisting
        a = _incline_accel("oracle", 0.3)
    assert abs(a - a_true) / a_true > 0.3, (a, a_true)


@pytest.mark.parametrize("backend", BACKENDS)
def test_incline_static_block_holds(backend):
    s = _incline(backend, 0.8)
    for _ in range(70):
        s.step(2e-4, 10)
    assert abs(float(s.v()[:, 0].mean())) < 0.02


# ---------------------------------------------------------------- 2. hydrostatic column
@pytest.mark.parametrize("backend", BACKENDS)
def test_hydrostatic_column_weight_on_scale(backend):
    dx = G32.dx
    h = dx / 2
    z0 = 4 * dx
    rho = 1000.0
    pos = _fill(0.14, 0.26, 0.14, 0.26, z0, z0 + 0.12, h)
    with _threaded(backend):
        s = _solver(backend, G32)
        s.load_particles(pos, np.full(len(pos), h ** 3, np.float32))
        s.set_material(newtonian(eta=0.5, density=rho, bulk_modulus=5.0e4), g=[0.0, 0.0, -GRAV])
        scale = s.add_box(center=(0.2, 0.2, z0 - 0.045), half_size=(0.12, 0.12, 0.045))
        for pt, nrm in (((0.14, 0, 0), (1, 0, 0)), ((0.26, 0, 0), (-1, 0, 0)),
                        ((0, 0.14, 0), (0, 1, 0)), ((0, 0.26, 0), (0, -1, 0))):
            s.add_plane(pt, nrm, "slip")
        weight = float(len(pos) * h ** 3 * rho * GRAV)
        for _ in range(100):
            s.step(2e-4, 10)
        _set_damping(s, 0.999)
        for _ in range(50):
            s.step(2e-4, 10)
        n = 60
        s.reset_tool_force(scale)
        for _ in range(n):
            s.step(2e-4, 10)
        fz = float(s.tool_force(scale, n * 2e-3)[2])
    assert fz < 0.0
    assert abs(-fz - weight) / weight < 0.08, (fz, weight)


# ---------------------------------------------------------------- 3. elastic bar
@pytest.mark.parametrize("backend", BACKENDS)
def test_elastic_bar_vibration_period(backend):
    grid = GridConfig(n_grid=48, grid_lim=0.4)
    h = grid.dx / 2
    E, rho, L = 1.0e6, 1000.0, 0.2
    x0 = 0.1
    pos = _fill(x0, x0 + L, 0.17, 0.22, 0.17, 0.22, h)
    s = _solver(backend, grid)
    s.load_particles(pos, np.full(len(pos), h ** 3, np.float32))
    s.set_material(elastic(E=E, nu=0.0, density=rho), g=[0.0, 0.0, 0.0])
    xc = x0 + L / 2
    v = np.zeros_like(pos)
    v[:, 0] = 0.1 * (pos[:, 0] - xc) / (L / 2)
    s.set_v(v)
    dt, sub = 5e-5, 10
    T = 2.0 * L / np.sqrt(E / rho)
    right = pos[:, 0] > xc
    times, dv = [], []
    for tick in range(int(3.2 * T / (dt * sub))):
        s.step(dt, sub)
        vx = s.v()[:, 0]
        times.append((tick + 1) * dt * sub)
        dv.append(float(vx[right].mean() - vx[~right].mean()))
    dv, times = np.asarray(dv), np.asarray(times)
    cross = np.where(np.diff(np.sign(dv)) != 0)[0]
    assert len(cross) >= 3
    t_cross = times[cross] - dv[cross] * (times[cross + 1] - times[cross]) / (dv[cross + 1] - dv[cross])
    T_meas = 2.0 * float(np.mean(np.diff(t_cross)))
    assert abs(T_meas - T) / T < 0.03, (T_meas, T)


# ---------------------------------------------------------------- 4. cup weighs its water
@pytest.mark.parametrize("backend", BACKENDS)
def test_static_cup_holds_water_and_reads_weight(backend):
    from mpm_amd.colliders.glass import GlassProfile, cavity_mask, cup_fill, solid_mask

    prof = GlassProfile()
    grid = GridConfig(n_grid=40, grid_lim=0.35)
    h = grid.dx / 2
    pos_local, vol = cup_fill(prof, h, fill_fraction=0.5)
    cup_pos = np.array([0.175, 0.175, 0.16])
    q_id = (1.0, 0.0, 0.0, 0.0)
    with _threaded(backend):
        s = _solver(backend, grid)
        s.load_particles((pos_local + cup_pos).astype(np.float32), vol)
        s.set_material(newtonian(eta=1.0e-3, density=1000.0, bulk_modulus=9.0e5))
        s.add_plane((0, 0, 3 * grid.dx), (0, 0, 1), "separate", friction=0.3)
        s.add_domain_walls()
        cup = s.add_cup(prof, cup_pos, q_id, friction=0.05)
        m_water = float(1000.0 * vol.sum())
        dt = 5.0e-5
        fz = []
        for _ in range(6):
            s.reset_cup_wrench(cup)
            s.step(dt, 500)
            fz.append(s.cup_wrench(cup, dt * 500)["force"][2])
            assert int(solid_mask(s.x(), cup_pos, q_id, prof).sum()) == 0, "wall penetration"
        x = s.x()
    assert cavity_mask(x, cup_pos, q_id, prof, pad=0.75 * h).all(), "water left the cup"
    f = float(np.mean(fz[-3:]))
    assert abs(f + GRAV * m_water) < 0.1 * GRAV * m_water, (f, -GRAV * m_water)


# ---------------------------------------------------------------- 5. mesh-SDF cup reaction
def _cup_sdf():
    """The reference's test cup (inner radius 40 mm, 8 mm wall, 90 mm tall, 10 mm base;
    test_sdf_collider.py:20-24) as an exact revolved SDF sampled on a 48^3 voxel grid with
    a 12 mm margin; body origin at the outside of the base."""
    from mpm_amd.colliders.glass import GlassProfile, glass_sdf_local
    from mpm_amd.geometry import sdf_from_function

    prof = GlassProfile(outer_radius=0.048, inner_radius=0.04, height=0.09, base_thickness=0.01,
                        fillet_radius=0.0)
    m = 0.012
    return sdf_from_function(lambda p: glass_sdf_local(p - np.array([0.0, 0.0, 0.045]), prof),
                             (-0.048 - m, -0.048 - m, -m), (0.048 + m, 0.048 + m, 0.09 + m), 48)


@pytest.mark.parametrize("backend", BACKENDS)
def test_sdf_cup_reaction_equals_contained_weight(backend):
    grid = GridConfig(n_grid=48, grid_lim=0.4)
    center = np.array([0.2, 0.2, 0.06])
    hh = grid.dx / 2
    pts = [(x, y, z) for x in np.arange(-0.03, 0.03, hh) for y in np.arange(-0.03, 0.03, hh)
           for z in np.arange(0.015, 0.05, hh) if x * x + y * y < 0.03 ** 2]
    pts = (np.array(pts) + center).astype(np.float32)
    vol = np.full(len(pts), hh ** 3, np.float32)
    s = _solver(backend, grid)
    s.load_particles(pts, vol)
    s.set_material(newtonian(eta=5.0, density=1000.0, bulk_modulus=5.0e5))
    hd = s.add_sdf_collider(_cup_sdf(), center=center, surface="separable", friction=0.3)
    dt = 2.0e-4
    fz = []
    for it in range(30):
        s.reset_sdf_force(hd)
        s.step(dt, 5)
        if it >= 15:
            fz.append(s.sdf_wrench(hd, dt * 5)["force"][2])
    weight = float(vol.sum() * 1000.0) * GRAV
    assert s.x()[:, 2].min() > center[2], "fluid leaked through the base"
    assert abs(abs(float(np.mean(fz))) - weight) < 0.25 * weight, (np.mean(fz), weight)


# ---------------------------------------------------------------- 6. domain walls
@pytest.mark.parametrize("backend", BACKENDS)
def test_block_thrown_at_the_walls_stays_in_and_slides(backend):
    grid = GridConfig(n_grid=32, grid_lim=1.0)
    h = grid.dx / 2
    pos = _fill(0.70, 0.82, 0.40, 0.52, 0.40, 0.52, h)
    s = _solver(backend, grid)
    s.load_particles(pos, np.full(len(pos), h ** 3, np.float32))
    s.set_material(elastic(E=2.0e4, nu=0.3, density=1000.0), g=[0.0, 0.0, 0.0])
    s.add_domain_walls()
    v = np.zeros_like(pos)
    v[:, 0], v[:, 1] = 2.0, 0.5  # into the +x wall, along +y
    s.set_v(v)
    for _ in range(30):
        s.step(2e-4, 10)
    x = s.x()
    assert x[:, 0].max() < 1.0 - 2 * grid.dx, "material passed the wall shell"
    assert float(s.v()[:, 1].mean()) > 0.4, "the wall held the tangential motion (sticky)"
