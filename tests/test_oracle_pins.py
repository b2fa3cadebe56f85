"""Pins the CPU oracle (the checker) against library routines, closed forms and
invariants, chosen so a dropped term, wrong sign, wrong index or transposed operand in
the restated reference kernels fails at least one check. No GPU needed."""
from __future__ import annotations

import numpy as np
import pytest
import scipy.linalg
import scipy.optimize

from oracle import sim as O

f32 = np.float32


def _single(pos, n_grid=16, lim=1.0, mat="jelly", **kw):
    s = O.OracleSim(np.asarray(pos, f32).reshape(-1, 3), np.full(len(np.atleast_2d(pos)), 1e-4, f32),
                    n_grid, lim)
    params = {"material": mat, "density": 1000.0, "E": 1e5, "nu": 0.3, "g": [0, 0, 0]}
    params.update(kw)
    s.set_parameters_dict(params)
    s.finalize_mu_lam()
    return s


# ------------------------------------------------------------------------------ SVD
@pytest.mark.parametrize("seed", range(6))
def test_svd_matches_numpy_and_convention(seed):
    rng = np.random.default_rng(seed)
    for k in range(200):
        A = rng.normal(size=(3, 3)).astype(f32)
        if k % 4 == 0:
            A = (np.eye(3) + 0.05 * rng.normal(size=(3, 3))).astype(f32)
        if k % 7 == 0:
            A[:, 2] = -A[:, 2]  # det < 0
        U, s, V = O.svd3(A)
        np.testing.assert_allclose(U @ np.diag(s) @ V.T, A, atol=5e-6 * np.abs(A).max() + 1e-7)
        np.testing.assert_allclose(U @ U.T, np.eye(3), atol=2e-6)
        np.testing.assert_allclose(V @ V.T, np.eye(3), atol=2e-6)
        assert np.linalg.det(U) > 0 and np.linalg.det(V) > 0
        ref = np.linalg.svd(A.astype(np.float64), compute_uv=False)
        np.testing.assert_allclose(np.abs(s), ref, rtol=1e-5, atol=1e-6)
        assert s[0] >= 0 and s[1] >= 0
        assert np.sign(s[2]) == np.sign(np.linalg.det(A.astype(np.float64))) or abs(s[2]) < 1e-6


def test_svd_degenerate_cases():
    for A in (np.zeros((3, 3)), np.eye(3), np.diag([2.0, 2.0, 2.0]), np.diag([1.0, 1.0, 0.0]),
              np.diag([3.0, 1.0, -2.0]), np.outer([1, 2, 3], [0.5, -1, 2])):
        A = A.astype(f32)
        U, s, V = O.svd3(A)
        assert np.isfinite(U).all() and np.isfinite(s).all() and np.isfinite(V).all()
        np.testing.assert_allclose(U @ np.diag(s) @ V.T, A, atol=1e-5)


# ------------------------------------------------------------------- B-spline transfers
@pytest.mark.parametrize("seed", range(4))
def test_p2g_conserves_mass_and_momentum(seed):
    """Sum of grid mass = particle mass; grid momentum = m v: the affine (C) and stress
    terms vanish by sum w (x_i - x_p) = 0 and sum grad w = 0."""
    rng = np.random.default_rng(seed)
    pos = rng.uniform(0.3, 0.7, size=(5, 3))
    s = _single(pos)
    s.v[:] = rng.normal(size=(5, 3))
    s.C[:] = rng.normal(size=(5, 3, 3)) * 10
    s.stress[:] = rng.normal(size=(5, 3, 3)) * 1e4
    lib = O.lib()
    lib.o_p2g(s._state(), s._model(), 1e-3)
    M = s.mass.astype(np.float64)
    np.testing.assert_allclose(s.grid_m.sum(), M.sum(), rtol=1e-6)
    np.testing.assert_allclose(s.grid_v_in.reshape(-1, 3).sum(0),
                               (M[:, None] * s.v).sum(0), rtol=1e-4, atol=1e-7)


def test_p2g_second_moment_and_single_node_weights():
    """sum w (x_i - x_p)(x_i - x_p)^T = dx^2/4 I for the quadratic B-spline (the APIC D
    matrix), and the affine term deposits m C (x_i - x_p) so the first moment of the
    deposited momentum is m C D."""
    pos = np.array([[0.513, 0.47, 0.52]])
    s = _single(pos)
    Cm = np.array([[1.0, 2.0, -1.0], [0.5, -3.0, 0.25], [2.0, 1.0, 1.0]], f32)
    s.C[0] = Cm
    O.lib().o_p2g(s._state(), s._model(), 0.0)
    dx = s.dx
    idx = np.argwhere(s.grid_m > 0)
    xi = idx * dx
    r = xi - pos[0]
    w = s.grid_m[tuple(idx.T)] / s.mass[0]
    np.testing.assert_allclose(w.sum(), 1.0, rtol=1e-6)
    np.testing.assert_allclose((w[:, None] * r).sum(0), 0.0, atol=1e-7)
    D = np.einsum("n,ni,nj->ij", w, r, r)
    np.testing.assert_allclose(D, np.eye(3) * dx * dx / 4, atol=1e-8)
    mom = s.grid_v_in[tuple(idx.T)]
    first = np.einsum("ni,nj->ij", mom, r)
    np.testing.assert_allclose(first, s.mass[0] * Cm @ D, rtol=2e-4, atol=1e-12)


def test_p2g_stress_term_is_minus_vol_tau_gradw():
    """With v = C = 0 the deposit is -dt vol tau grad w, grad w taken w.r.t. the particle
    position; since sum grad w (x_i - x_p)^T = I its first moment is -dt vol tau."""
    pos = np.array([[0.49, 0.505, 0.53]])
    s = _single(pos)
    tau = np.array([[3.0, 1.0, 0.0], [1.0, -2.0, 0.5], [0.0, 0.5, 1.0]], f32) * 1e3
    s.stress[0] = tau
    dt = 1e-3
    O.lib().o_p2g(s._state(), s._model(), dt)
    idx = np.argwhere(s.grid_m > 0)
    r = idx * s.dx - pos[0]
    mom = s.grid_v_in[tuple(idx.T)].astype(np.float64)
    first = np.einsum("ni,nj->ij", mom, r)
    np.testing.assert_allclose(first, -dt * s.vol[0] * tau, rtol=1e-4, atol=1e-10)


def test_g2p_reproduces_affine_field():
    """An affine grid velocity v(x) = a + B x is reproduced exactly: v_p = a + B x_p,
    C = B and L = B (APIC/MLS exactness)."""
    rng = np.random.default_rng(3)
    pos = rng.uniform(0.35, 0.65, size=(6, 3))
    s = _single(pos, n_grid=16)
    a = np.array([0.1, -0.2, 0.3])
    B = rng.normal(size=(3, 3))
    g = np.stack(np.meshgrid(*[np.arange(16) * s.dx] * 3, indexing="ij"), -1)
    s.grid_v_out[:] = (a + g @ B.T).astype(f32)
    x0 = s.x.copy()
    dt = 1e-4
    O.lib().o_g2p(s._state(), s._model(), dt)
    vp = a + x0 @ B.T
    np.testing.assert_allclose(s.v, vp, rtol=1e-4, atol=1e-5)
    np.testing.assert_allclose(s.C, np.broadcast_to(B, (6, 3, 3)), rtol=1e-3, atol=2e-4)
    np.testing.assert_allclose(s.L, np.broadcast_to(B, (6, 3, 3)), rtol=1e-3, atol=2e-4)
    np.testing.assert_allclose(s.x, x0 + dt * s.v, atol=1e-7)
    np.testing.assert_allclose(s.F_trial, np.broadcast_to(np.eye(3) + dt * s.L[0], (6, 3, 3)),
                               atol=1e-5)


# ------------------------------------------------------------------------ materials
def _stress_of(Ft, mat, **kw):
    s = _single([[0.5, 0.5, 0.5]], mat=mat, **kw)
    s.F_trial[0] = np.asarray(Ft, f32)
    O.lib().o_stress_particle(s._state(), s._model(), kw.pop("dt", 1e-4), 0)
    return s


def test_fcr_stress_matches_polar_closed_form():
    rng = np.random.default_rng(0)
    for _ in range(20):
        F = (np.eye(3) + 0.2 * rng.normal(size=(3, 3))).astype(f32)
        s = _stress_of(F, "jelly")
        mu, lam = float(s.mu[0]), float(s.lam[0])
        R, _ = scipy.linalg.polar(F.astype(np.float64))
        J = np.linalg.det(F.astype(np.float64))
        tau = 2 * mu * (F - R) @ F.T + lam * J * (J - 1) * np.eye(3)
        np.testing.assert_allclose(s.stress[0], 0.5 * (tau + tau.T), rtol=2e-4, atol=1e-2)
        np.testing.assert_array_equal(s.F[0], F)  # jelly: no return map


def test_mu_lam_from_E_nu():
    s = _single([[0.5, 0.5, 0.5]], E=1e5, nu=0.3)
    np.testing.assert_allclose(s.mu[0], 1e5 / 2.6, rtol=1e-6)
    np.testing.assert_allclose(s.lam[0], 1e5 * 0.3 / (1.3 * 0.4), rtol=1e-6)


def test_von_mises_diagonal_return_and_hardening():
    """Diagonal F: yield iff |dev tau| > yield; the returned log-strain sits on the yield
    surface |dev(2 mu eps)| = yield (+ the 1e-6 regulariser), volumetric part kept; with
    hardening the yield stress grows by 2 mu xi dgamma."""
    E, nu, ys = 2e5, 0.3, 1e3
    F = np.diag([1.05, 0.97, 0.99]).astype(f32)
    s = _stress_of(F, "metal", E=E, nu=nu, yield_stress=ys, hardening=1.0, xi=0.5)
    mu = E / (2 * (1 + nu))
    eps = np.log(np.diag(F).astype(np.float64))
    dev = eps - eps.mean()
    dg = np.linalg.norm(dev) + 1e-6 - ys / (2 * mu)
    eps_new = eps - dg / (np.linalg.norm(dev) + 1e-6) * dev
    Fe = s.F[0].astype(np.float64)
    np.testing.assert_allclose(np.diag(Fe), np.exp(eps_new), rtol=2e-5)
    np.testing.assert_allclose(Fe - np.diag(np.diag(Fe)), 0, atol=1e-6)
    np.testing.assert_allclose(s.yield_stress[0], ys + 2 * mu * 0.5 * dg, rtol=1e-4)
    # Hencky stress of the returned F: diag(2 mu eps + lam tr eps)
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    tau = 2 * mu * eps_new + lam * eps_new.sum()
    np.testing.assert_allclose(np.diag(s.stress[0]), tau, rtol=1e-3)
    # below yield: untouched
    s2 = _stress_of(np.diag([1.001, 1.0, 1.0]), "metal", E=E, nu=nu, yield_stress=ys)
    np.testing.assert_array_equal(s2.F[0], np.diag([1.001, 1.0, 1.0]).astype(f32))


def test_mu_i_bisection_matches_brentq():
    """mu(I) plastic correction: gdp solves tau_bar - G dt gdp - mu(I) p = 0 with
    I = gdp d sqrt(rho_s / p_C) (48 bisection steps)."""
    E, nu, dt = 1e6, 0.3, 1e-4
    mu_s, dmu, I0, d, rho = 0.38, 0.26, 0.28, 1e-3, 2650.0
    F = np.diag([1.02, 0.97, 0.99]).astype(f32)
    F[0, 1] = 0.03
    s = _stress_of(F, "mu_i_sand", E=E, nu=nu, mu_s=mu_s, delta_mu=dmu, I0=I0,
                   grain_diameter=d, grain_density=rho, dt=dt)
    G = E / (2 * (1 + nu))
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    U, sig, Vt = np.linalg.svd(F.astype(np.float64))
    eps = np.log(sig)
    tr = eps.sum()
    assert tr < 0
    dev = eps - tr / 3
    pK = -(2 * G / 3 + lam) * tr
    tb = 2 * G * np.linalg.norm(dev) / np.sqrt(2)
    pC = pK / np.prod(sig)
    Ic = d * np.sqrt(rho / pC)
    assert tb > mu_s * pK
    g = lambda x: tb - G * dt * x - (mu_s + dmu * x * Ic / (x * Ic + I0)) * pK  # noqa: E731
    gd = scipy.optimize.brentq(g, 0.0, tb / (G * dt), xtol=1e-14)
    scale = (tb - G * dt * gd) / tb
    e_new = dev * scale + tr / 3
    Fe = U @ np.diag(np.exp(e_new)) @ Vt
    np.testing.assert_allclose(s.F[0], Fe, rtol=2e-4, atol=2e-6)


def test_newtonian_closed_form():
    """mat 10: F -> J^(1/3) I; tau = J(-p I + 2 eta_app dev D), p = -K (J^-1.1 - 1),
    eta_app = eta + tau_y / gd + K_hb gd^(n-1), gd = sqrt(2 dev D : dev D + 0.02^2)."""
    K, eta, ty, kh, nh = 1e4, 10.0, 200.0, 50.0, 0.5
    F = np.diag([1.01, 0.99, 0.995]).astype(f32)
    s = _single([[0.5, 0.5, 0.5]], mat="newtonian", bulk_modulus=K, plastic_viscosity=eta,
                yield_stress=ty, hardening=kh, softening=nh)
    L = np.array([[0.5, 2.0, 0.0], [-1.0, 0.1, 0.3], [0.2, 0.0, -0.4]], f32)
    s.F_trial[0] = F
    s.L[0] = L
    O.lib().o_stress_particle(s._state(), s._model(), 1e-4, 0)
    J = np.linalg.det(F.astype(np.float64))
    np.testing.assert_allclose(s.F[0], np.eye(3) * J ** (1 / 3), rtol=1e-6)
    p = -K * (J ** -1.1 - 1)
    D = 0.5 * (L + L.T).astype(np.float64)
    Dd = D - np.trace(D) / 3 * np.eye(3)
    gd = np.sqrt(2 * (Dd * Dd).sum() + 0.02 ** 2)
    ea = eta + ty / gd + kh * gd ** (nh - 1)
    tau = J * (p * np.eye(3) + 2 * ea * Dd)
    np.testing.assert_allclose(s.stress[0], tau, rtol=2e-5, atol=1e-3)


def test_tabulated_viscous_equals_parametric():
    """A table sampling the parametric eta_app(gd) reproduces mat 10 to interpolation
    error (reference test_tabulated_material's equivalence)."""
    eta, ty = 5.0, 20.0
    smin, smax, n = -2.0, 3.0, 2001
    sg = np.linspace(smin, smax, n)
    tab = eta + ty / 10 ** sg
    L = np.array([[0.5, 2.0, 0.0], [-1.0, 0.1, 0.3], [0.2, 0.0, -0.4]], f32)
    F = np.diag([1.01, 0.99, 0.995]).astype(f32)
    a = _single([[0.5] * 3], mat="newtonian", bulk_modulus=1e4, plastic_viscosity=eta, yield_stress=ty)
    b = _single([[0.5] * 3], mat="tabulated_viscous", bulk_modulus=1e4, eta_table=tab,
                eta_table_smin=smin, eta_table_smax=smax)
    for s in (a, b):
        s.F_trial[0] = F
        s.L[0] = L
        O.lib().o_stress_particle(s._state(), s._model(), 1e-4, 0)
    np.testing.assert_allclose(b.stress[0], a.stress[0], rtol=5e-3, atol=1e-2)


def test_water_and_stationary():
    F = np.diag([1.02, 1.0, 0.97]).astype(f32)
    s = _stress_of(F, "fluid", bulk_modulus=2e4)
    J = float(np.linalg.det(F.astype(np.float64)))
    p = -2e4 * (J ** -1.1 - 1)
    np.testing.assert_allclose(s.stress[0], np.eye(3) * p * J, rtol=1e-4)
    s = _stress_of(F, "stationary")
    np.testing.assert_array_equal(s.F[0], np.eye(3, dtype=f32))
    np.testing.assert_array_equal(s.stress[0], np.zeros((3, 3), f32))


# ------------------------------------------------------------------------ grid ops
def test_normalize_threshold_and_gravity():
    s = _single([[0.5, 0.5, 0.5]], g=[0.0, -9.8, 0.0])
    s.grid_m[1, 2, 3] = 2.0
    s.grid_v_in[1, 2, 3] = [4.0, 2.0, -2.0]
    s.grid_m[2, 2, 2] = 1e-16     # below the 1e-15 threshold: v_out untouched
    s.grid_v_in[2, 2, 2] = [1.0, 1.0, 1.0]
    s.grid_v_out[2, 2, 2] = [7.0, 7.0, 7.0]
    lo, dims = O.i3((0, 0, 0)), O.i3((16, 16, 16))
    O.lib().o_grid_normalize(s._state(), s._model(), 0.01, lo, dims)
    np.testing.assert_allclose(s.grid_v_out[1, 2, 3], [2.0, 1.0 - 0.098, -1.0], rtol=1e-6)
    np.testing.assert_array_equal(s.grid_v_out[2, 2, 2], [7.0, 7.0, 7.0])


def test_box_collider_impulse_closed_form():
    s = _single([[0.5, 0.5, 0.5]])
    s.set_velocity_on_cuboid((0.5, 0.5, 0.5), (0.1, 0.1, 0.1), (0.0, 0.0, -0.05))
    s.grid_m[8, 8, 8] = 2.0
    s.grid_v_out[8, 8, 8] = [0.1, 0.2, 0.3]
    s.grid_m[0, 0, 0] = 1.0          # outside the box
    s.grid_v_out[0, 0, 0] = [1.0, 1.0, 1.0]
    s._apply_grid_bc(1e-3)
    p = s.colliders[0][1]
    np.testing.assert_allclose(p._force, 2.0 * (np.array([0.1, 0.2, 0.3]) - [0, 0, -0.05]), rtol=1e-6)
    np.testing.assert_allclose(s.grid_v_out[8, 8, 8], [0, 0, -0.05])
    np.testing.assert_array_equal(s.grid_v_out[0, 0, 0], [1, 1, 1])
    # modify: the box centre advances dt * v after the launch
    np.testing.assert_allclose(p.point[2], 0.5 - 0.05e-3, rtol=1e-6)


def test_cup_sdf_matches_host_twin():
    """The revolved-glass SDF inside the cup kernel agrees with the host twin (the
    reference keeps the two in lockstep): nodes touched by the kernel are exactly those
    with twin sdf < contact_band."""
    import mpm_amd.colliders.glass as glass

    prof = glass.GlassProfile(outer_radius=0.05, inner_radius=0.04, height=0.12,
                              base_thickness=0.02, fillet_radius=0.006)
    n = 48
    s = _single([[0.5, 0.5, 0.5]], n_grid=n, lim=0.24)
    center = np.array([0.12, 0.12, 0.12])
    band = 0.5 * s.dx
    s.add_revolved_sdf_collider(center, np.eye(3), outer_radius=prof.outer_radius,
                                inner_radius=prof.inner_radius, half_height=prof.half_height,
                                inner_floor_z=prof.inner_floor_z, fillet_radius=prof.fillet_radius,
                                friction=0.0, sticky_depth=0.9 * s.dx, contact_band=band)
    s.grid_m[:] = 1.0
    s.grid_v_out[:] = np.array([0.3, -0.2, 0.1], f32)
    s._apply_grid_bc(1e-4)
    touched = np.any(s.grid_v_out != np.array([0.3, -0.2, 0.1], f32), axis=-1)
    g = np.stack(np.meshgrid(*[np.arange(n) * s.dx] * 3, indexing="ij"), -1)
    sdf = glass.glass_sdf_local(g - center, prof)
    deep = sdf < -0.9 * s.dx - 1e-6
    # deep nodes are grabbed (velocity -> 0 for a static cup)
    assert touched[deep].all()
    # nodes clearly outside the band are never touched
    assert not touched[sdf > band + 1e-6].any()
