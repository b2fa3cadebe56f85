"""Hand-written adjoint (mpm_amd.adjoint.Rollout) against central finite differences of
the CPU oracle's forward rollout (GPU).

The adjoint differentiates the discrete substep map; the oracle evaluates the same map
sequentially in fp32, so central differences of oracle rollouts (eps = 1e-3 relative, as
stated for the squeeze identification) pin the gradient without trusting any part of
the GPU code. Scene: the squeeze workload at reduced size (32^3 grid, 16^3 particles,
von Mises E = 2e5, nu = 0.3, yield 1e3 between two closing box plates on a sticky floor),
so a rollout costs the oracle well under a second."""
from __future__ import annotations

import numpy as np
import pytest

from mpm_amd import Solver
from mpm_amd.adjoint import Rollout
from mpm_amd.scenes import SqueezeScene
from oracle.sim import OracleSolver

pytestmark = pytest.mark.gpu

DT = 1e-4


def _scene():
    return SqueezeScene(n_grid=32, per_side=16, dt=DT)


def _oracle(sc, mu, lam, ys, v0, x0=None, C0=None, Ft0=None):
    o = OracleSolver(sc.grid)
    o.load_particles(sc.pos if x0 is None else x0, sc.vol)
    o.set_material(sc.material)
    o.add_plane((0, 0, sc.floor), (0, 0, 1), "sticky")
    plates = [o.add_box(c, sc.plate_half, v) for c, v in zip(sc.plate_centers, sc.plate_velocity)]
    o._sim.mu[:] = mu
    o._sim.lam[:] = lam
    o._sim.yield_stress[:] = ys
    o._sim.v[:] = v0
    if C0 is not None:
        o._sim.C[:] = C0
    if Ft0 is not None:
        o._sim.F_trial[:] = Ft0
    return o, plates


def _oracle_rollout(sc, substeps, mu, lam, ys, v0, **kw):
    o, plates = _oracle(sc, mu, lam, ys, v0, **kw)
    o.step(DT, substeps)
    imp = np.stack([o.tool_force(h, 1.0) for h in plates])
    return o.x().astype(np.float64), imp.astype(np.float64)


def _lame(E, nu):
    return E / (2.0 * (1.0 + nu)), E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu))


def test_rollout_gradients_match_central_differences():
    """60 substeps with checkpoints every 20: d loss / d (E, nu, yield) and the
    directional derivative along a random initial-velocity perturbation, against
    central differences of the oracle rollout at eps = 1e-3 relative."""
    import torch

    sc = _scene()
    s = sc.build(Solver(grid=sc.grid, device="hip"))
    n = s.n_particles
    S = 60
    ro = Rollout(s, DT, S, checkpoint_every=20)
    dev = s._sim.device
    E0, nu0, y0 = 2.0e5, 0.3, 1.0e3
    # seeded target rollout: yield x 1.1
    with torch.no_grad():
        xt, wt = ro.run(torch.tensor(E0, dtype=torch.float64, device=dev),
                        torch.tensor(nu0, dtype=torch.float64, device=dev),
                        torch.tensor(1.1 * y0, dtype=torch.float64, device=dev),
                        torch.zeros(n, 3, device=dev))
    x_star, w_star = xt.double(), wt[:, :3].double()
    E = torch.tensor(E0, dtype=torch.float64, device=dev, requires_grad=True)
    nu = torch.tensor(nu0, dtype=torch.float64, device=dev, requires_grad=True)
    ys = torch.tensor(y0, dtype=torch.float64, device=dev, requires_grad=True)
    v0 = torch.zeros(n, 3, device=dev, requires_grad=True)
    xT, w = ro.run(E, nu, ys, v0)
    loss = ((xT.double() - x_star) ** 2).sum() + ((w[:, :3].double() - w_star) ** 2).sum()
    loss.backward()
    g = {"E": E.grad.item(), "nu": nu.grad.item(), "yield": ys.grad.item()}

    xs, ws = x_star.cpu().numpy(), w_star.cpu().numpy()

    def oracle_loss(E_, nu_, y_, v_):
        mu, lam = _lame(E_, nu_)
        x, imp = _oracle_rollout(sc, S, np.float32(mu), np.float32(lam), np.float32(y_), v_)
        return ((x - xs) ** 2).sum() + ((imp - ws) ** 2).sum()

    zero_v = np.zeros((n, 3), np.float32)
    fd = {}
    for name, (a, b, c) in {"E": (E0, nu0, y0), "nu": (E0, nu0, y0), "yield": (E0, nu0, y0)}.items():
        base = {"E": a, "nu": b, "yield": c}[name]
        h = 1e-3 * base
        args_p = [a, b, c]
        args_m = [a, b, c]
        i = ["E", "nu", "yield"].index(name)
        args_p[i] += h
        args_m[i] -= h
        fd[name] = (oracle_loss(*args_p, zero_v) - oracle_loss(*args_m, zero_v)) / (2 * h)
    rng = np.random.default_rng(3)
    d = rng.choice([-1.0, 1.0], size=(n, 3)).astype(np.float32)
    hv = 1e-3
    fd_v = (oracle_loss(E0, nu0, y0, hv * d) - oracle_loss(E0, nu0, y0, -hv * d)) / (2 * hv)
    ad_v = float((v0.grad.double().cpu().numpy() * d).sum())
    report = {k: (g[k], fd[k]) for k in g}
    report["v0.d"] = (ad_v, fd_v)
    print(report)
    for k, (a, f) in report.items():
        assert abs(a - f) <= 1e-3 * abs(f) + 1e-12, f"{k}: adjoint {a:.6e} vs central difference {f:.6e}"


def test_one_substep_directional_derivatives():
    """One substep from a random deformed state (some particles past yield, random v
    and affine C): directional derivatives of a random linear functional of (x_T,
    plate impulses) along x0, v0, mu, lam and yield perturbations."""
    import torch

    sc = _scene()
    s = sc.build(Solver(grid=sc.grid, device="hip"))
    n = s.n_particles
    rng = np.random.default_rng(5)
    v0 = rng.uniform(-0.2, 0.2, (n, 3)).astype(np.float32)
    C0 = rng.uniform(-2.0, 2.0, (n, 3, 3)).astype(np.float32)
    # deviatoric strain up to 2 %: particles on both sides of the yield surface
    Ft0 = (np.eye(3) + 0.02 * rng.uniform(0, 1, (n, 1, 1)) * rng.uniform(-1, 1, (n, 3, 3))).astype(np.float32)
    sim = s._sim
    sim.C.copy_(torch.from_numpy(C0.reshape(n, 9).T.copy()))
    sim.F_trial.copy_(torch.from_numpy(Ft0.reshape(n, 9).T.copy()))
    ro = Rollout(s, DT, 1, checkpoint_every=1)
    dev = sim.device
    mu0, lam0 = (np.float32(v) for v in _lame(2.0e5, 0.3))
    y0 = np.float32(1.0e3)
    cx = rng.standard_normal((n, 3))
    cw = rng.standard_normal((2, 3)) * 1e3
    mu = torch.full((n,), float(mu0), device=dev, requires_grad=True)
    lam = torch.full((n,), float(lam0), device=dev, requires_grad=True)
    ys = torch.full((n,), float(y0), device=dev, requires_grad=True)
    v = torch.from_numpy(v0).to(dev).requires_grad_(True)
    x0 = torch.from_numpy(sc.pos.copy()).to(dev).requires_grad_(True)
    xT, w = ro(mu, lam, ys, v, x0)
    loss = (xT.double() * torch.from_numpy(cx).to(dev)).sum() + \
        (w[:, :3].double() * torch.from_numpy(cw).to(dev)).sum()
    loss.backward()

    def oracle_loss(mu_, lam_, y_, v_, x_):
        x, imp = _oracle_rollout(sc, 1, mu_, lam_, y_, v_, x0=x_, C0=C0, Ft0=Ft0)
        return (x * cx).sum() + (imp * cw).sum()

    dx = sc.grid.dx
    dirs = {
        "x0": (x0, (rng.choice([-1.0, 1.0], (n, 3)) * 1e-3 * dx).astype(np.float32)),
        "v0": (v, (rng.choice([-1.0, 1.0], (n, 3)) * 1e-3).astype(np.float32)),
        "mu": (mu, (rng.uniform(0.5, 1.5, n) * 1e-3 * mu0).astype(np.float32)),
        "lam": (lam, (rng.uniform(0.5, 1.5, n) * 1e-3 * lam0).astype(np.float32)),
        "yield": (ys, (rng.uniform(0.5, 1.5, n) * 1e-3 * y0).astype(np.float32)),
    }
    base = {"mu": np.full(n, mu0, np.float32), "lam": np.full(n, lam0, np.float32),
            "yield": np.full(n, y0, np.float32), "v0": v0, "x0": sc.pos.astype(np.float32)}
    for name, (t, d) in dirs.items():
        ad = float((t.grad.double().cpu().numpy() * d).sum())
        args_p = dict(base)
        args_m = dict(base)
        args_p[name] = base[name] + d
        args_m[name] = base[name] - d
        lp = oracle_loss(args_p["mu"], args_p["lam"], args_p["yield"], args_p["v0"], args_p["x0"])
        lm = oracle_loss(args_m["mu"], args_m["lam"], args_m["yield"], args_m["v0"], args_m["x0"])
        fd = (lp - lm) / 2.0
        print(name, ad, fd)
        assert abs(ad - fd) <= 2e-3 * abs(fd) + 1e-9, f"{name}: adjoint {ad:.6e} vs central difference {fd:.6e}"


def test_unsupported_material_is_refused():
    from mpm_amd.materials import newtonian

    sc = _scene()
    s = Solver(grid=sc.grid, device="hip")
    s.load_particles(sc.pos, sc.vol)
    s.set_material(newtonian(eta=1.0))
    with pytest.raises(NotImplementedError):
        Rollout(s, DT, 10)
