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
@pytest.mark.parametrize("k,label,group", [(13, "mu_i", "sigma_min"), (13, "sand", "sigma_min"),
                                             (14, "jelly", "det<0"), (14, "mu_i", "det<0")])
def test_fragile_spectra_bar_has_teeth(k, label, group):
    """The bar of the GPU's fragile-spectrum SVD case (per spectrum group, within 10x the
    oracle's spread under a 1-ulp perturbation of F_trial) fails on the two mistakes a
    fp32 SVD plausibly makes: singular values read off the eigenvalues of a float32
    A^T A (planted mistake 13: the conditioning squared, ~3 % at sigma_min = 1e-3), and
    the reflection of a det F < 0 state carried by the largest axis (planted mistake 14:
    a valid decomposition, but another R = U V^T and other clamped logs)."""
    from test_hip_parity import MATERIALS, _fragile_start, _mat_run

    material = dict(MATERIALS)[label]
    F0, groups = _fragile_start(3000)
    rng = np.random.default_rng(99)
    F1 = (F0 * (1 + rng.choice([-1, 1], F0.shape) * 2.0 ** -23)).astype(np.float32)
    kw = dict(ticks=(1,), dt=1.0e-7)
    ora = _mat_run(O.OracleSolver, material, F0=F0, **kw).stress().reshape(-1, 9).astype(np.float64)
    ulp = _mat_run(O.OracleSolver, material, F0=F1, **kw).stress().reshape(-1, 9).astype(np.float64)
    with O.mutant(k):
        mut = _mat_run(O.OracleSolver, material, F0=F0, **kw).stress().reshape(-1, 9).astype(np.float64)
    sl = groups[group]
    spread = np.abs(ulp[sl] - ora[sl]).max()
    err = np.abs(mut[sl] - ora[sl]).max()
    bar = 10.0 * spread + 1e-6 * np.nanmax(np.abs(ora))
    print(f"mutant {k} {label} {group}: {err:.3e} vs bar {bar:.3e}")
    assert err > 10.0 * bar, f"mutant {k}: {err:.3e} vs bar {bar:.3e}"


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


# ------------------------------------------- sand, plasticine, mu(I)-Phi(I), tabulated mu(I)
def _rot(rng):
    q = np.linalg.qr(rng.normal(size=(3, 3)))[0]
    return q * np.sign(np.linalg.det(q))


def _frame(rng, e):
    """F = R1 diag(exp(e)) R2^T in float32 with random proper rotations."""
    R1, R2 = _rot(rng), _rot(rng)
    return (R1 @ np.diag(np.exp(e)) @ R2.T).astype(f32), R1, R2


def _run(F, mat, Jp=None, dt=1e-4, **kw):
    s = _single([[0.5, 0.5, 0.5]], mat=mat, **kw)
    s.F_trial[0] = np.asarray(F, f32)
    if Jp is not None:
        s.Jp[0] = Jp
    O.lib().o_stress_particle(s._state(), s._model(), dt, 0)
    return s


def _logsv(F):
    return np.log(np.linalg.svd(np.asarray(F, np.float64), compute_uv=False))


def test_sand_cone_projection_closed_form():
    """Drucker-Prager return (mat 2, sand_return_mapping mpm_utils.py:543-580) with
    alpha = sqrt(2/3) 2 sin(phi) / (3 - sin(phi)) from the friction angle: inside the cone
    (|dev e| + c tr e <= 0, c = (3 lam + 2 mu) / (2 mu) alpha) F is untouched; outside it
    in tension F -> U V^T (scipy's polar rotation); outside it in compression the
    log-strain is projected onto the cone along its deviatoric part: trace kept,
    |dev e_new| = -c tr e, deviatoric direction and principal frame kept."""
    E, nu, phi = 1e5, 0.3, 30.0
    mu, lam = E / 2.6, E * 0.3 / (1.3 * 0.4)
    sp = np.sin(np.radians(phi))
    c = (3 * lam + 2 * mu) / (2 * mu) * np.sqrt(2.0 / 3.0) * 2 * sp / (3 - sp)
    rng = np.random.default_rng(31)
    for kind in ("inside", "outside", "tension"):
        for _ in range(5):
            d = rng.normal(size=3)
            d -= d.mean()
            d /= np.linalg.norm(d)
            tr = 0.01 if kind == "tension" else -0.03
            mag = {"inside": 0.5, "outside": 2.0, "tension": 1.0}[kind] * c * 0.03
            e = tr / 3 + mag * d
            F, R1, R2 = _frame(rng, e)
            s = _run(F, "sand", E=E, nu=nu, friction_angle=phi)
            Fn = s.F[0].astype(np.float64)
            if kind == "inside":
                np.testing.assert_array_equal(s.F[0], F)
            elif kind == "tension":
                R, _ = scipy.linalg.polar(F.astype(np.float64))
                np.testing.assert_allclose(Fn, R, atol=3e-6)
            else:
                e_new = tr / 3 + (-c * tr) * d
                np.testing.assert_allclose(Fn, R1 @ np.diag(np.exp(e_new)) @ R2.T, atol=3e-6)
                ln = np.sort(_logsv(Fn))
                np.testing.assert_allclose(ln.sum(), tr, atol=2e-6)
                np.testing.assert_allclose(np.linalg.norm(ln - ln.mean()), -c * tr, rtol=2e-4)


def test_drucker_prager_stress_equals_hencky():
    """kirchoff_stress_drucker_prager (mpm_utils.py:109-124), U diag((2 mu log s + lam
    sum log s) / s) V^T F^T, is the Hencky stress U diag(2 mu log s + lam sum log s) U^T
    (V^T F^T = diag(s) U^T); checked on the stress of the returned F for states inside the
    cone, where F is the trial F."""
    E, nu = 1e5, 0.3
    mu, lam = E / 2.6, E * 0.3 / (1.3 * 0.4)
    rng = np.random.default_rng(32)
    for _ in range(10):
        e = -0.02 + 0.002 * rng.normal(size=3)
        F, _, _ = _frame(rng, e)
        s = _run(F, "sand", E=E, nu=nu, friction_angle=40.0)
        U, sv, _ = np.linalg.svd(s.F[0].astype(np.float64))
        ls = np.log(sv)
        tau = U @ np.diag(2 * mu * ls + lam * ls.sum()) @ U.T
        np.testing.assert_allclose(s.stress[0], tau, rtol=1e-4, atol=1e-4 * np.abs(tau).max())


def test_plasticine_softening_and_damage():
    """von_mises_return_mapping_with_damage (mat 5, mpm_utils.py:173-225) with FCR stress:
    a yielding state returns like von Mises and the yield stress drops by softening x
    |plastic log-strain increment| = softening x dg |h| / (|h| + 1e-6); a large softening
    drives it to <= 0, which zeroes mu and lam (so the FCR stress vanishes) while F still
    takes the return computed with the old yield stress; a particle already at yield
    stress <= 0 keeps its trial F."""
    E, nu, ys, soft = 2e5, 0.3, 1e3, 5e3
    mu, lam = E / 2.6, E * 0.3 / (1.3 * 0.4)
    F = np.diag([1.05, 0.97, 0.99]).astype(f32)
    eps = np.log(np.diag(F).astype(np.float64))
    h = eps - eps.mean()
    hn = np.linalg.norm(h) + 1e-6
    dg = hn - ys / (2 * mu)
    e_new = eps - dg / hn * h
    s = _run(F, "plasticine", E=E, nu=nu, yield_stress=ys, softening=soft)
    np.testing.assert_allclose(np.diag(s.F[0]), np.exp(e_new), rtol=2e-5)
    np.testing.assert_allclose(s.yield_stress[0], ys - soft * dg * np.linalg.norm(h) / hn, rtol=1e-4)
    Fe = np.diag(np.exp(e_new))
    J = np.prod(np.exp(e_new))
    tau = 2 * mu * (Fe - np.eye(3)) @ Fe.T + lam * J * (J - 1) * np.eye(3)  # FCR, R = I
    np.testing.assert_allclose(s.stress[0], tau, rtol=1e-3, atol=1e-2)
    # damage to zero
    d = _run(F, "plasticine", E=E, nu=nu, yield_stress=ys, softening=1e9)
    assert d.yield_stress[0] <= 0 and d.mu[0] == 0 and d.lam[0] == 0
    np.testing.assert_allclose(np.diag(d.F[0]), np.exp(e_new), rtol=2e-5)
    np.testing.assert_array_equal(d.stress[0], np.zeros((3, 3), f32))
    # fully damaged already: the trial F is kept
    z = _run(F, "plasticine", E=E, nu=nu, yield_stress=0.0, softening=soft)
    np.testing.assert_array_equal(z.F[0], F)


def _phi_setup(E=1e6, nu=0.3):
    return dict(E=E, nu=nu, mu_s=0.38, delta_mu=0.26, I0=0.28, grain_diameter=1e-3, grain_density=2650.0,
                phi_chi=0.2)


def test_mu_i_phi_pressure_branches_and_reference_volume():
    """mu_i_phi_return_mapping (mat 11, mpm_utils.py:434-522) with the stress of
    :525-540: compaction pressure p = max(K (Jp_ref / J - 1), 0), K = E / (3 (1 - 2 nu)),
    Jp_ref = Jp or 1 when Jp < 0.5. No pressure (J >= Jp_ref): F -> exp(mean e) U V^T and
    zero stress. Pressure below the mu_s cone: F kept, Jp = Jp_ref, tau = U diag(2 G dev e
    - p J) U^T. A stored Jp of 0.8 (>= 0.5) is the reference volume, so a 1 % compression
    is still dilated (no pressure)."""
    kw = _phi_setup()
    E, nu = kw["E"], kw["nu"]
    G = E / (2 * (1 + nu))
    K = E / (3 * (1 - 2 * nu))
    rng = np.random.default_rng(33)
    # expansion: isotropic return, zero stress
    F, R1, R2 = _frame(rng, np.array([0.012, -0.002, 0.0]))
    s = _run(F, "mu_i_phi", **kw)
    e = _logsv(F)
    np.testing.assert_allclose(s.F[0], np.exp(e.mean()) * (R1 @ R2.T), atol=3e-6)
    assert s.Jp[0] == 1.0
    np.testing.assert_allclose(s.stress[0], 0.0, atol=1e-2)
    # compression below the cone
    e = np.array([-0.004, -0.006, -0.005])
    F, R1, _ = _frame(rng, e)
    s = _run(F, "mu_i_phi", **kw)
    np.testing.assert_array_equal(s.F[0], F)
    J = np.exp(e.sum())
    p = K * (1.0 / J - 1.0)
    assert np.sqrt(2) * G * np.linalg.norm(e - e.mean()) < 0.38 * p * J  # below the cone
    tau = R1 @ np.diag(2 * G * (e - e.mean()) - p * J) @ R1.T
    np.testing.assert_allclose(s.stress[0], tau, rtol=2e-4, atol=2e-4 * np.abs(tau).max())
    np.testing.assert_allclose(s.Jp[0], 1.0)
    # stored reference volume 0.8: a 1 % compression has no pressure
    F, R1, R2 = _frame(rng, np.array([-0.004, -0.003, -0.003]))
    s = _run(F, "mu_i_phi", Jp=0.8, **kw)
    np.testing.assert_allclose(s.F[0], np.exp(_logsv(F).mean()) * (R1 @ R2.T), atol=3e-6)
    np.testing.assert_allclose(s.Jp[0], 0.8)


def test_mu_i_phi_yield_root_and_dilatancy():
    """Yielding mu(I)-Phi(I) state: gdp solves tau_bar - G dt gdp - mu(I) p_K = 0 with
    p_K = p J, I = gdp d sqrt(rho_s / p) (scipy brentq on the same g), the deviatoric
    log-strain is scaled by (tau_bar - G dt gdp) / tau_bar, the reference volume relaxes
    as Jp = Jp_ref + min(8 gdp dt, 1) (j_cs - Jp_ref) with j_cs = 1 / max(1 - chi I /
    (I + I0), 1e-3), and the stress uses the returned F and the updated Jp."""
    kw = _phi_setup()
    E, nu, dt = kw["E"], kw["nu"], 1e-4
    G = E / (2 * (1 + nu))
    K = E / (3 * (1 - 2 * nu))
    rng = np.random.default_rng(34)
    e = np.array([0.02, -0.035, -0.015])
    F, R1, R2 = _frame(rng, e)
    s = _run(F, "mu_i_phi", dt=dt, **kw)
    e = _logsv(F)[[0, 1, 2]]
    # frame-free checks on singular values (sorted)
    J = np.exp(e.sum())
    p = K * (1.0 / J - 1.0)
    pK = p * J
    tb = np.sqrt(2) * G * np.linalg.norm(e - e.mean())
    assert p > 0 and tb > 0.38 * pK
    Ic = 1e-3 * np.sqrt(2650.0 / p)
    g = lambda x: tb - G * dt * x - (0.38 + 0.26 * x * Ic / (x * Ic + 0.28)) * pK  # noqa: E731
    gd = scipy.optimize.brentq(g, 0.0, tb / (G * dt), xtol=1e-14)
    scale = (tb - G * dt * gd) / tb
    e_new = np.sort(e.mean() + scale * (e - e.mean()))
    np.testing.assert_allclose(np.sort(_logsv(s.F[0])), e_new, atol=3e-6)
    I = gd * Ic
    jcs = 1.0 / max(1.0 - 0.2 * I / (I + 0.28), 1e-3)
    Jp = 1.0 + min(8.0 * gd * dt, 1.0) * (jcs - 1.0)
    np.testing.assert_allclose(s.Jp[0], Jp, rtol=1e-5)
    # stress of the returned F with the new reference volume
    U, sv, _ = np.linalg.svd(s.F[0].astype(np.float64))
    ln = np.log(sv)
    Jn = np.prod(sv)
    pn = max(K * (Jp / Jn - 1.0), 0.0)
    tau = U @ np.diag(2 * G * (ln - ln.mean()) - pn * Jn) @ U.T
    np.testing.assert_allclose(s.stress[0], tau, rtol=2e-4, atol=2e-4 * np.abs(tau).max())


def test_tabulated_mu_i_equals_parametric_and_uses_table_head_as_threshold():
    """mu_i_tabulated_return_mapping (mat 13, mpm_utils.py:328-407): with a table sampling
    the parametric mu(I) on log10 I in [-6, 1] the return equals the parametric mu(I)
    return (mat 9) to interpolation error; the yield threshold is the table's first entry
    (a table starting at 5 keeps a state the parametric law yields)."""
    E, nu, dt = 1e6, 0.3, 1e-4
    smin, smax, n = -6.0, 1.0, 4001
    sg = np.linspace(smin, smax, n)
    I = 10.0 ** sg
    tab = 0.38 + 0.26 * I / (I + 0.28)
    rng = np.random.default_rng(35)
    e = np.array([0.02, -0.035, -0.015])
    F, _, _ = _frame(rng, e)
    common = dict(E=E, nu=nu, grain_diameter=1e-3, grain_density=2650.0, dt=dt)
    a = _run(F, "mu_i_sand", mu_s=0.38, delta_mu=0.26, I0=0.28, **common)
    b = _run(F, "tabulated_mu_i", eta_table=tab, eta_table_smin=smin, eta_table_smax=smax, **common)
    assert np.abs(a.F[0] - F).max() > 1e-3  # the parametric law yields
    np.testing.assert_allclose(b.F[0], a.F[0], atol=2e-6)
    np.testing.assert_allclose(b.stress[0], a.stress[0], rtol=1e-3, atol=1e-3 * np.abs(a.stress[0]).max())
    high = tab.copy()
    high[0] = 5.0
    c = _run(F, "tabulated_mu_i", eta_table=high, eta_table_smin=smin, eta_table_smax=smax, **common)
    np.testing.assert_array_equal(c.F[0], F)


# ------------------------------------------------------------------------ rigid bodies
def _rigid_cube(n_grid=32, z0=0.5, fused=False, planes=()):
    from mpm_amd.core.solver import GridConfig

    g = GridConfig(n_grid=n_grid, grid_lim=1.0)
    h = g.dx / 2
    ax = np.arange(0.4 + h / 2, 0.6, h)
    pts = np.stack(np.meshgrid(ax, ax, ax - 0.4 + z0, indexing="ij"), -1).reshape(-1, 3).astype(f32)
    o = O.OracleSolver(g, fused=fused)
    o.load_particles(pts, np.full(len(pts), h ** 3, f32))
    o.set_material("rigid", density=1000.0)
    o.set_material_range(0, len(pts), "rigid", obj_id=0)
    for pl in planes:
        o.add_plane(*pl)
    o.finalize_rigid_bodies()
    return o, pts, h


def test_rigid_finalize_mass_com_and_lattice_inertia():
    """finalize_rigid_bodies (mpm_solver_warp.py:834-873): mass = sum m, COM = mass-weighted
    mean, inertia of an n^3 lattice of spacing h = M h^2 (n^2 - 1) / 6 on the diagonal
    (2 x the variance of n equally spaced points) and zero off it; x_ref = x - COM."""
    o, pts, h = _rigid_cube()
    s = o._sim
    M = float(s.mass.sum(dtype=np.float64))
    np.testing.assert_allclose(s.rigid_mass[0], M, rtol=1e-6)
    # (float32 sums over the body, as the reference's numpy finalize)
    np.testing.assert_allclose(s.rigid_x_cm[0], pts.astype(np.float64).mean(0), atol=1e-5)
    n = round(len(pts) ** (1 / 3))
    inertia = np.linalg.inv(s.rigid_inv_inertia_body[0].astype(np.float64))
    np.testing.assert_allclose(inertia, M * h * h * (n * n - 1) / 6 * np.eye(3), rtol=1e-4, atol=1e-9)
    np.testing.assert_allclose(s.x_ref, pts - s.rigid_x_cm[0], atol=1e-7)


def test_rigid_free_fall_and_orientation():
    """A rigid cube in free fall (no colliders): v_cm = g t, COM falls g t (t + dt) / 2
    (the symplectic update, x += v dt after v), the orientation stays orthonormal and
    every particle moves with the body."""
    o, pts, _ = _rigid_cube(z0=0.6)
    dt, n = 1e-4, 60
    z0 = float(o._sim.rigid_x_cm[0][2])
    o.step(dt, n)
    st = o.rigid_state(0)
    t = n * dt
    np.testing.assert_allclose(st["v"], [0.0, 0.0, -9.81 * t], atol=5e-6)
    np.testing.assert_allclose(st["com"][2], z0 - 9.81 * dt * dt * n * (n + 1) / 2, atol=2e-6)
    np.testing.assert_allclose(st["R"] @ st["R"].T, np.eye(3), atol=1e-6)
    dz = o.x()[:, 2] - pts[:, 2]
    assert np.ptp(dz) < 1e-5 and abs(dz.mean() - (st["com"][2] - z0)) < 2e-6


def test_rigid_restitution_reverses_contact_normal_velocity():
    """The restitution impulse (mpm_solver_warp.py:874-962): at the deepest particle the
    normal contact velocity after the impulse is -e times the one before (the impulse is
    -(1 + e) v_n over the effective inverse mass 1/M + ((I^-1 (r x n)) x r) . n), the body
    is pushed out of the plane, and without friction the impulse is J n at the contact:
    the COM velocity changes along n only and the angular impulse is r x (M dv_cm), i.e.
    I_world d(omega) = r x M d(v_cm). Checked on the restitution step alone, from a
    tilted, spinning body just through the plane."""
    o, pts, _ = _rigid_cube(z0=0.5, planes=[((0, 0, 0.3), (0, 0, 1), "separate", 0.0, 0.5)])
    s = o._sim
    th = 0.3
    Rz = np.array([[1, 0, 0], [0, np.cos(th), -np.sin(th)], [0, np.sin(th), np.cos(th)]], f32)
    s.rigid_orientation[0] = Rz
    s.rigid_x_cm[0] = [0.5, 0.5, 0.3 + 0.08]
    s.rigid_v_cm[0] = [0.2, -0.1, -1.5]
    s.rigid_omega[0] = [0.5, 2.0, -1.0]
    x_world = s.rigid_x_cm[0] + (Rz @ s.x_ref.T).T
    d = (x_world - np.array([0, 0, 0.3])) @ np.array([0, 0, 1.0])
    i = int(np.argmin(d))
    assert d[i] < 0
    r = (Rz @ s.x_ref[i]).astype(np.float64)
    v0, w0 = s.rigid_v_cm[0].astype(np.float64), s.rigid_omega[0].astype(np.float64)
    v_before = v0 + np.cross(w0, r)
    s._apply_rigid_restitution()
    v1, w1 = s.rigid_v_cm[0].astype(np.float64), s.rigid_omega[0].astype(np.float64)
    v_after = v1 + np.cross(w1, r)
    np.testing.assert_allclose(v_after[2], -0.5 * v_before[2], rtol=1e-5)
    np.testing.assert_allclose(v1[:2], v0[:2], atol=1e-7)
    M = float(s.rigid_mass[0])
    I_world = Rz.astype(np.float64) @ np.linalg.inv(s.rigid_inv_inertia_body[0].astype(np.float64)) @ Rz.T
    np.testing.assert_allclose(I_world @ (w1 - w0), np.cross(r, M * (v1 - v0)), rtol=1e-4,
                               atol=1e-6 * np.abs(M * (v1 - v0)).max())
    x_new = s.rigid_x_cm[0] + (Rz @ s.x_ref.T).T
    assert ((x_new - [0, 0, 0.3]) @ [0, 0, 1.0]).min() > -1e-6


# ------------------------------------------------------------------ particle modifiers
def _modifier_block(fused=False):
    from mpm_amd.core.solver import GridConfig

    g = GridConfig(n_grid=32, grid_lim=1.0)
    h = g.dx / 2
    ax = np.arange(0.3 + h / 2, 0.7, h)
    pts = np.stack(np.meshgrid(ax, ax, np.arange(0.3 + h / 2, 0.5, h), indexing="ij"), -1).reshape(-1, 3).astype(f32)
    o = O.OracleSolver(g, fused=fused)
    o.load_particles(pts, np.full(len(pts), h ** 3, f32))
    o.set_material("jelly", E=1e5, nu=0.3, density=1000.0, g=[0.0, 0.0, 0.0])
    return o, pts


def test_modifier_selection_and_closed_forms():
    """Particle modifiers (mpm_solver_warp.py:2385-2539): the box selection is strict
    |x - point| < size per axis; the cylinder selection |axial| < half height and radial
    < radius; on the first substep, an active translation sets v = velocity exactly; a
    rotation sets v = -s_r n x r + s_t n for the radial offset r (speed s_r |r|, normal to
    r; the reference's second axis is axis1 x n, so a positive rotation_scale turns about
    -n); an impulse adds force / mass * dt, only inside its window; unselected particles
    are untouched."""
    o, pts = _modifier_block()
    s = o._sim
    s.enforce_particle_velocity_translation((0.35, 0.5, 0.4), (0.04, 0.3, 0.3), (0.2, 0.0, 0.0), 0.0, 1.0)
    s.enforce_particle_velocity_rotation((0.6, 0.5, 0.4), (0, 0, 1), (0.3, 0.08), 3.0, 0.1, 0.0, 1.0)
    s.add_impulse_on_particles((0.0, 0.0, 5e-4), 1e-4, point=(0.5, 0.35, 0.4), size=(0.05, 0.04, 0.3), num_dt=3)
    box = np.all(np.abs(pts - [0.35, 0.5, 0.4]) < [0.04, 0.3, 0.3], axis=1)
    off = pts - [0.6, 0.5, 0.4]
    cyl = (np.abs(off[:, 2]) < 0.3) & (np.linalg.norm(off[:, :2], axis=1) < 0.08)
    imp = np.all(np.abs(pts - [0.5, 0.35, 0.4]) < [0.05, 0.04, 0.3], axis=1) & ~box & ~cyl
    np.testing.assert_array_equal(s.particle_velocity_modifier_params[0]["mask"], box)
    np.testing.assert_array_equal(s.particle_velocity_modifier_params[1]["mask"], cyl)
    s._apply_particle_modifiers(1e-4)
    np.testing.assert_array_equal(s.v[box], np.tile([0.2, 0.0, 0.0], (box.sum(), 1)).astype(f32))
    w_expected = -3.0 * np.stack([-off[cyl, 1], off[cyl, 0], np.zeros(cyl.sum())], 1) + [0, 0, 0.1]
    np.testing.assert_allclose(s.v[cyl], w_expected, atol=2e-6)
    dv = 5e-4 / s.mass[imp] * 1e-4
    np.testing.assert_allclose(s.v[imp][:, 2], dv, rtol=1e-6)
    assert np.all(s.v[~(box | cyl | imp)] == 0)
    s.time = 3.5e-4  # past the impulse window (3 dt), the overrides still active
    before = s.v.copy()
    s._apply_particle_modifiers(1e-4)
    np.testing.assert_array_equal(s.v[imp], before[imp])
