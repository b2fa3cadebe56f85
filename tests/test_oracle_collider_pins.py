"""Per-node pins of the oracle's collider responses, grid damping and covariance transport
(the checker). Each check states a physical law rather than the kernel's arithmetic:

* a separable (Coulomb) contact removes the approaching normal velocity relative to the
  surface and takes at most mu |v_n| of the tangential speed, all of it while sliding,
  without turning the tangential direction; separating or outside nodes stay untouched;
* slip removes only the normal relative velocity; sticky moves with the surface;
* the reaction wrench is the momentum removed, sum m (v_in - v_new), with torque
  sum r x m (v_in - v_new) about the collider's centre;
* domain walls only stop motion out of the domain (the reference's bounding box is not
  sticky), within three nodes of each face;
* the point-cloud collider stops exactly the nodes within L1 distance `padding` of a
  point's cell (the reference's six-neighbour dilation, repeated);
* damping scales the grid velocity by `scale`;
* the covariance update is one explicit Euler step of dS/dt = L S + S L^T, so it agrees
  with the exact transport expm(L dt) S expm(L dt)^T to O(dt^2).

Every check then runs on an oracle built with one planted mistake (oracle.sim.mutant:
a flipped friction or normal sign, a swapped cross product, a transposed L, ...) and
must fail there."""
from __future__ import annotations

import numpy as np
import pytest
import scipy.linalg

from mpm_amd import GridConfig
from oracle import sim as O

f32 = np.float32
DT = 1e-4


def _sim(n=16, lim=1.0):
    s = O.OracleSim(np.full((1, 3), 0.5, f32), np.full(1, 1e-4, f32), n, lim)
    s.set_parameters_dict({"material": "jelly", "density": 1000.0, "E": 1e5, "nu": 0.3, "g": [0, 0, 0]})
    return s


def _nodes(s):
    n = s.n_grid
    return np.stack(np.meshgrid(*[np.arange(n) * s.dx] * 3, indexing="ij"), -1).astype(np.float64)


def _load_grid(s, rng, frac_massive=0.7):
    n = s.n_grid
    m = rng.uniform(0.5, 2.0, (n, n, n)) * (rng.uniform(size=(n, n, n)) < frac_massive)
    s.grid_m[:] = m.astype(f32)
    s.grid_v_out[:] = rng.normal(size=(n, n, n, 3)).astype(f32)
    return s.grid_m.astype(np.float64).copy(), s.grid_v_out.astype(np.float64).copy()


def _unit(v):
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


def _coulomb_ok(v0, v1, vs, nrm, mu, tol):
    """True where v1 is the Coulomb response of v0 against a surface moving at vs with
    unit normal nrm (all (k, 3) float64): the approaching normal relative velocity is
    gone, the tangential relative speed dropped by min(|v_t|, mu |v_n|) along its own
    direction; separating relative velocities are kept."""
    r0, r1 = v0 - vs, v1 - vs
    vn = (r0 * nrm).sum(-1)
    t0 = r0 - vn[:, None] * nrm
    t0l = np.linalg.norm(t0, axis=-1)
    app = vn < 0
    ok = np.ones(len(v0), bool)
    ok[~app] = np.abs(r1[~app] - r0[~app]).max(-1) <= tol * (1 + np.abs(r0[~app]).max(-1))
    n1 = (r1[app] * nrm[app]).sum(-1)
    t1 = r1[app] - n1[:, None] * nrm[app]
    want = np.maximum(t0l[app] - mu * np.abs(vn[app]), 0.0)
    dir_ok = (t1 * t0[app]).sum(-1) >= -tol  # not turned around
    ok[app] = (np.abs(n1) <= tol * (1 + np.abs(vn[app]))) & \
        (np.abs(np.linalg.norm(t1, axis=-1) - want) <= tol * (1 + t0l[app])) & dir_ok & \
        (np.linalg.norm(np.cross(t1, t0[app]), axis=-1) <= tol * (1 + t0l[app]) ** 2)
    return ok


# ---------------------------------------------------------------------------- plane
def _plane_check(surface, mu, seed=0):
    """Run the oracle's plane collider on a random grid; return the fraction of nodes
    whose response breaks the law for `surface`."""
    rng = np.random.default_rng(seed)
    s = _sim()
    nrm = _unit(rng.normal(size=3))
    pt = rng.uniform(0.35, 0.65, 3)
    s.add_surface_collider(pt, nrm, surface, friction=mu)
    m, v0 = _load_grid(s, rng)
    s.time = 0.0
    s._apply_grid_bc(DT)
    v1 = s.grid_v_out.astype(np.float64)
    nw = np.asarray([s.colliders[0][1].normal[i] for i in range(3)], np.float64)
    side = ((_nodes(s) - pt) @ nw) < 0
    act = side & (m > 0)
    bad = np.zeros(m.shape, bool)
    bad[~act] = np.any(v1[~act] != v0[~act], axis=-1)  # untouched, bitwise
    a, b = v0[act], v1[act]
    if surface == "sticky":
        bad[act] = np.abs(b).max(-1) != 0.0
    elif surface == "slip":
        vn = a @ nw
        bad[act] = np.abs(b - (a - vn[:, None] * nw)).max(-1) > 1e-5
    else:
        bad[act] = ~_coulomb_ok(a, b, np.zeros_like(a), np.broadcast_to(nw, a.shape), mu, 2e-5)
    assert act.sum() > 200 and (~act & (m > 0)).sum() > 200  # both sides populated
    return bad.mean()


@pytest.mark.parametrize("surface,mu", [("sticky", 0.0), ("slip", 0.0), ("separate", 0.0),
                                        ("separate", 0.3), ("separate", 2.0)])
def test_plane_response_obeys_contact_law(surface, mu):
    for seed in range(3):
        assert _plane_check(surface, mu, seed) == 0.0


def test_plane_pins_have_teeth():
    with O.mutant(1):  # friction sign flipped: sliding nodes speed up
        assert _plane_check("separate", 0.3) > 0.01
    with O.mutant(2):  # side test flipped: the wrong half-space is constrained
        assert _plane_check("slip", 0.0) > 0.1


def test_plane_time_window():
    rng = np.random.default_rng(1)
    s = _sim()
    s.add_surface_collider((0.5, 0.5, 0.5), (0, 0, 1), "sticky", start_time=0.1, end_time=0.2)
    m, v0 = _load_grid(s, rng)
    s.time = 0.05
    s._apply_grid_bc(DT)
    np.testing.assert_array_equal(s.grid_v_out, v0.astype(f32))
    s.time = 0.15
    s._apply_grid_bc(DT)
    assert np.all(s.grid_v_out[:, :, :7][m[:, :, :7] > 0] == 0.0)


# ---------------------------------------------------------------------------- walls
def _walls_check():
    rng = np.random.default_rng(4)
    s = _sim()
    s.add_bounding_box()
    m, v0 = _load_grid(s, rng)
    s.time = 0.0
    s._apply_grid_bc(DT)
    v1 = s.grid_v_out.astype(np.float64)
    n = s.n_grid
    idx = np.stack(np.meshgrid(*[np.arange(n)] * 3, indexing="ij"), -1)
    want = v0.copy()
    for a in range(3):
        lo = (idx[..., a] < 3) & (v0[..., a] < 0) & (m > 0)
        hi = (idx[..., a] >= n - 3) & (v0[..., a] > 0) & (m > 0)
        want[..., a][lo | hi] = 0.0  # only the outward component, never the others
    return np.mean(np.any(v1 != want, axis=-1))


def test_domain_walls_stop_only_outward_motion():
    assert _walls_check() == 0.0


def test_walls_pin_has_teeth():
    with O.mutant(3):  # the low x shell stops inward instead of outward motion
        assert _walls_check() > 0.01


# ------------------------------------------------------------------------ mesh SDF
def _sphere_field(R, res, cell):
    """Exact sphere SDF |x| - R and its gradient x / |x| on a body-frame voxel grid."""
    origin = -0.5 * (res - 1) * cell * np.ones(3)
    g = origin + cell * np.stack(np.meshgrid(*[np.arange(res)] * 3, indexing="ij"), -1)
    r = np.linalg.norm(g, axis=-1)
    grad = g / np.maximum(r, 1e-12)[..., None]
    return (r - R).astype(f32), grad.astype(f32), origin


def _sdf_check(surface, mu=0.4, seed=0, wrench=False):
    rng = np.random.default_rng(seed)
    s = _sim(n=32)
    R, res, cell = 0.17, 48, 0.0125
    val, grad, origin = _sphere_field(R, res, cell)
    c = np.array([0.5, 0.49, 0.51])
    q = _unit(rng.normal(size=4))  # xyzw: a rotated sphere is still a sphere
    u, w = rng.normal(size=3) * 0.2, rng.normal(size=3)
    band = s.dx
    s.add_sdf_collider(val, grad, origin, cell, c, quat=q, velocity=u, omega=w, band=band,
                       surface=surface, friction=mu)
    m, v0 = _load_grid(s, rng, frac_massive=0.9)
    s.time = 0.0
    p = s.colliders[0][1]
    f0 = np.array(p._force, np.float64)
    t0 = np.array(p._torque, np.float64)
    s._apply_grid_bc(DT)
    v1 = s.grid_v_out.astype(np.float64)
    rel = _nodes(s) - c
    d = np.linalg.norm(rel, axis=-1) - R
    # surely within the band (trilerp error < 2e-3), and 3 cells off the centre, where
    # the gradient field x / |x| is singular and its trilerp no unit normal
    inside = (d <= band - 2e-3) & (d > 3 * cell - R) & (m > 0)
    outside = (d > band + 2e-3) | (m == 0)
    nr = _unit(rel)
    vs = u + np.cross(w, rel)
    bad_out = np.any(v1[outside] != v0[outside], axis=-1).mean()
    if wrench:
        hit = np.any(v1 != v0, axis=-1)
        imp = m[..., None] * (v0 - v1)
        F = imp[hit].sum(0)
        T = np.cross(rel[hit], imp[hit]).sum(0)
        fe = np.abs(np.array(p._force) - f0 - F).max() / np.abs(F).max()
        te = np.abs(np.array(p._torque) - t0 - T).max() / np.abs(T).max()
        return fe, te
    a, b, vsi, ni = v0[inside], v1[inside], vs[inside], nr[inside]
    if surface == "sticky":
        bad_in = np.abs(b - vsi).max(-1) > 1e-5
    elif surface == "slip":
        r0, r1 = a - vsi, b - vsi
        vn = (r0 * ni).sum(-1)
        # normal from the trilerped gradient: within ~1e-3 rad of the radial direction
        bad_in = (np.abs((r1 * ni).sum(-1)) > 5e-3 * (1 + np.abs(r0).max(-1))) | \
            (np.abs(r1 - (r0 - vn[:, None] * ni)).max(-1) > 5e-3 * (1 + np.abs(r0).max(-1)))
    else:
        # skip nodes whose approach is within the normal's error of grazing
        vn = ((a - vsi) * ni).sum(-1)
        clear = np.abs(vn) > 2e-2 * np.linalg.norm(a - vsi, axis=-1)
        bad_in = ~_coulomb_ok(a[clear], b[clear], vsi[clear], ni[clear], mu, 5e-3)
    assert inside.sum() > 500
    return max(bad_out, np.mean(bad_in))


@pytest.mark.parametrize("surface", ["sticky", "slip", "separable"])
def test_sdf_collider_response_on_an_analytic_sphere(surface):
    for seed in range(2):
        assert _sdf_check(surface, seed=seed) == 0.0


def test_sdf_wrench_is_the_removed_momentum():
    fe, te = _sdf_check("separable", wrench=True)
    assert fe < 1e-5 and te < 1e-5


def test_sdf_pins_have_teeth():
    with O.mutant(4):  # normal sign flipped
        assert _sdf_check("separable") > 0.05
    with O.mutant(5):  # friction sign flipped
        assert _sdf_check("separable") > 0.01
    with O.mutant(12):  # slip keeps the normal velocity
        assert _sdf_check("slip") > 0.05


# ------------------------------------------------------------------ revolved cup
def _cup_setup(rng, mu, omega=(0.0, 0.0, 0.0)):
    import mpm_amd.colliders.glass as glass

    # a wall 2.9 cells thick, so nodes deeper than the sticky depth exist
    prof = glass.GlassProfile(outer_radius=0.14, inner_radius=0.08, height=0.3,
                              base_thickness=0.04, fillet_radius=0.01)
    s = _sim(n=48)
    c = np.array([0.5, 0.5, 0.5])
    s.add_revolved_sdf_collider(c, np.eye(3), velocity=(0.01, -0.02, 0.03), omega=omega,
                                outer_radius=prof.outer_radius, inner_radius=prof.inner_radius,
                                half_height=prof.half_height, inner_floor_z=prof.inner_floor_z,
                                fillet_radius=prof.fillet_radius, friction=mu,
                                sticky_depth=0.9 * s.dx, contact_band=0.5 * s.dx)
    m, v0 = _load_grid(s, rng, frac_massive=0.9)
    return s, prof, c, m, v0


def _cup_check(mu=0.3, seed=0, wrench=False):
    """Nodes at mid height, away from rim, floor and fillets, where the glass SDF is
    max(r - R_out, R_in - r): the outer wall's normal points radially out, the cavity
    wall's radially in."""
    rng = np.random.default_rng(seed)
    w = (0.0, 0.0, 1.5)
    s, prof, c, m, v0 = _cup_setup(rng, mu, omega=w)
    s.time = 0.0
    p = s.colliders[0][1]
    s._apply_grid_bc(DT)
    v1 = s.grid_v_out.astype(np.float64)
    rel = _nodes(s) - c
    if wrench:
        hit = np.any(v1 != v0, axis=-1)
        imp = m[..., None] * (v0 - v1)
        F, T = imp[hit].sum(0), np.cross(rel[hit], imp[hit]).sum(0)
        return (np.abs(np.array(p._force) - F).max() / np.abs(F).max(),
                np.abs(np.array(p._torque) - T).max() / np.abs(T).max())
    r = np.hypot(rel[..., 0], rel[..., 1])
    mid = (np.abs(rel[..., 2]) < 0.5 * prof.half_height) & (m > 0) & (r > 0.03)
    dx = s.dx
    d_out, d_in = r - prof.outer_radius, prof.inner_radius - r
    sdf = np.maximum(d_out, d_in)
    vs = np.asarray((0.01, -0.02, 0.03)) + np.cross(np.asarray(w), rel)
    radial = _unit(np.concatenate([rel[..., :2], np.zeros_like(rel[..., :1])], -1))
    nrm = np.where((d_out >= d_in)[..., None], radial, -radial)
    far = mid & (sdf > 0.5 * dx + 1e-4)
    deep = mid & (sdf < -0.9 * dx - 1e-4)
    shallow = mid & (sdf > -0.9 * dx + 1e-4) & (sdf < 0.5 * dx - 1e-4)
    assert far.sum() > 100 and deep.sum() > 20 and shallow.sum() > 100
    bad = [np.any(v1[far] != v0[far], axis=-1),               # beyond the band: untouched
           np.abs(v1[deep] - vs[deep]).max(-1) > 1e-5]         # grabbed by the solid
    a, b, vv, nn = v0[shallow], v1[shallow], vs[shallow], nrm[shallow]
    vn = ((a - vv) * nn).sum(-1)
    clear = np.abs(vn) > 2e-2 * np.linalg.norm(a - vv, axis=-1)
    bad.append(~_coulomb_ok(a[clear], b[clear], vv[clear], nn[clear], mu, 5e-3))
    return max(float(np.mean(x)) for x in bad)


def test_cup_grab_and_coulomb_branches():
    for seed in range(2):
        assert _cup_check(seed=seed) == 0.0


def test_cup_wrench_is_the_removed_momentum():
    fe, te = _cup_check(wrench=True)
    assert fe < 1e-5 and te < 1e-5


def test_cup_pins_have_teeth():
    with O.mutant(6):  # Coulomb friction sign flipped
        assert _cup_check() > 0.01
    with O.mutant(7):  # grab / Coulomb depth test flipped
        assert _cup_check() > 0.05
    with O.mutant(11):  # torque as imp x r
        assert _cup_check(wrench=True)[1] > 0.5


# -------------------------------------------------------------------- point cloud
def _pointcloud_check(padding):
    rng = np.random.default_rng(7)
    s = _sim(n=24)
    pts = rng.uniform(0.2, 0.8, size=(40, 3))
    s.add_point_cloud_collider(pts, padding=padding)
    m, v0 = _load_grid(s, rng)
    s.time = 0.0
    s._apply_grid_bc(DT)
    n = s.n_grid
    cells = np.clip((pts * s.inv_dx).astype(np.int32), 0, n - 1)
    idx = np.stack(np.meshgrid(*[np.arange(n)] * 3, indexing="ij"), -1)
    l1 = np.abs(idx[:, :, :, None, :] - cells[None, None, None]).sum(-1).min(-1)
    occ = l1 <= padding                     # brute force: L1 ball of radius `padding`
    want = v0.copy()
    want[occ & (m > 0)] = 0.0
    return np.mean(np.any(s.grid_v_out.astype(np.float64) != want, axis=-1))


@pytest.mark.parametrize("padding", [0, 1, 2])
def test_pointcloud_stops_exactly_the_dilated_cells(padding):
    assert _pointcloud_check(padding) == 0.0


def test_pointcloud_pin_has_teeth():
    with O.mutant(8):
        assert _pointcloud_check(1) > 0.1


# ------------------------------------------------------------------------ damping
def _damping_ratio(scale):
    rng = np.random.default_rng(9)
    s = _sim()
    s.grid_v_out[:] = rng.normal(size=s.grid_v_out.shape).astype(f32)
    v0 = s.grid_v_out.copy()
    O.lib().o_grid_damping(s._state(), s._model(), scale, O.i3((0, 0, 0)), O.i3((16, 16, 16)))
    return s.grid_v_out / v0


def test_damping_scales_grid_velocity():
    np.testing.assert_allclose(_damping_ratio(0.999), 0.999, rtol=2e-7)
    with O.mutant(9):
        assert np.abs(_damping_ratio(0.999) - 0.999).max() > 1e-3


def test_damping_applies_only_below_one_in_the_substep():
    """add_damping_via_grid runs only when grid_v_damping_scale < 1 (the reference's
    default 1.1 means off): a free particle keeps its speed at 1.1 and loses a factor
    0.99 per substep at 0.99 (the gather of a uniformly scaled field scales v)."""
    out = {}
    for scale in (1.1, 0.99):
        s = O.OracleSolver(GridConfig(16, 1.0))
        s.load_particles(np.array([[0.5, 0.5, 0.5]], f32), np.array([1e-4], f32))
        s.set_material("jelly", E=1e3, nu=0.3, density=1000.0, g=[0.0, 0.0, 0.0],
                       grid_v_damping_scale=scale)
        s.set_v(np.array([[0.2, 0.0, 0.0]], f32))
        s.step(DT, 5)
        out[scale] = float(s.v()[0, 0])
    assert out[1.1] == pytest.approx(0.2, rel=1e-5)
    assert out[0.99] == pytest.approx(0.2 * 0.99 ** 5, rel=1e-4)


# --------------------------------------------------------------- covariance transport
def _cov_error(seed=0):
    rng = np.random.default_rng(seed)
    pos = rng.uniform(0.35, 0.65, size=(6, 3))
    s = O.OracleSim(pos.astype(f32), np.full(6, 1e-4, f32), 16, 1.0)
    s.set_parameters_dict({"material": "jelly", "density": 1000.0, "E": 1e5, "nu": 0.3, "g": [0, 0, 0]})
    A = rng.normal(size=(6, 3, 3)) * 1e-3
    S0 = A @ A.transpose(0, 2, 1) + 1e-6 * np.eye(3)
    iu = ([0, 0, 0, 1, 1, 2], [0, 1, 2, 1, 2, 2])
    s.cov[:] = S0[:, iu[0], iu[1]].reshape(-1).astype(f32)
    s.update_cov_with_F = 1
    B = rng.normal(size=(3, 3)) * 20.0  # a non-symmetric velocity gradient
    g = np.stack(np.meshgrid(*[np.arange(16) * s.dx] * 3, indexing="ij"), -1)
    s.grid_v_out[:] = (g @ B.T).astype(f32)
    dt = 2e-4
    O.lib().o_g2p(s._state(), s._model(), dt)
    E = scipy.linalg.expm(B * dt)
    exact = E @ S0 @ E.T
    got = np.zeros((6, 3, 3))
    c = s.cov.reshape(6, 6).astype(np.float64)
    for k, (i, j) in enumerate(zip(*iu)):
        got[:, i, j] = got[:, j, i] = c[:, k]
    step = S0 + dt * (B @ S0 + S0 @ B.T)
    return (np.abs(got - exact).max() / np.abs(exact).max(),
            np.abs(got - step).max() / np.abs(step).max())


def test_cov_update_is_an_euler_step_of_the_exact_transport():
    err_exact, err_step = _cov_error()
    # |B dt| ~ 0.012: the Euler step differs from expm by O((B dt)^2) ~ 1e-4, a
    # transposed L by O(|B - B^T| dt) ~ 1e-2
    assert err_exact < 1e-3
    assert err_step < 2e-4  # and matches the step itself to the gather's fp32 accuracy
    with O.mutant(10):  # L transposed
        assert _cov_error()[0] > 3e-3
