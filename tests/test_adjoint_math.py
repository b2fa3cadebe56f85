"""The constitutive adjoints of the reverse substep (csrc/mpm_adjoint_math.h: fixed-corotated
elasticity, von Mises, local mu(I), generalized-Newtonian fluid) against float64 central
differences of the forward maps, on the CPU.

The header is compiled for the host (tests/host/adjoint_math_check.hip; no HIP call, no
GPU) and fed random states; each forward map is restated here in float64 from the
reference (mpm_utils.py: kirchoff_stress_FCR :8-15, von_mises_return_mapping :128-170, mu_i_return_mapping
:247-324, kirchoff_stress_newtonian :28-55 with stress_update_particle's fluid branch
:1057-1149, Hencky stress :228-243). The checked quantity is the adjoint of
J = <F, Fb> + <tau, tb> with respect to F_trial, L and every parameter, for states on both
sides of each yield surface (kept 2 % away from it), in tension, with repeated or nearly
repeated singular values (the divided-difference guard), and, for fixed-corotated
elasticity, large stretches and an inverted F (a negative singular value). The GPU tests check the
same kernels end to end against differences of whole rollouts; this test pins the
formulas to float32 rounding (F_trial and L adjoints within 2e-4 of the largest entry,
parameter derivatives within 1e-3 of the stress term they scale)."""
from __future__ import annotations

import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
SQ2 = np.sqrt(2.0)


# ---------------------------------------------------------------- float64 forward maps
def _svd(F):
    U, s, Vt = np.linalg.svd(F)
    return U, s, Vt.T


def _hencky(F, mu, lam):
    U, s, _ = _svd(F)
    e = np.log(np.maximum(s, 1e-6))
    t = U @ np.diag(2 * mu * e + lam * e.sum()) @ U.T
    return 0.5 * (t + t.T)


def _fcr(F, mu, lam):
    # wp.svd3's decomposition has rotations U and V and gives the smallest singular value
    # the sign of det F, so R = U V^T is a rotation; numpy's keeps s >= 0 and puts the sign
    # into U V^T, which the last column's flip takes back out
    U, s, V = _svd(F)
    R = U @ np.diag([1.0, 1.0, np.linalg.det(U @ V.T)]) @ V.T
    J = np.linalg.det(F)
    t = 2 * mu * (F - R) @ F.T + lam * J * (J - 1) * np.eye(3)
    return F, 0.5 * (t + t.T)


def _sand(Ft, mu, lam, alpha):
    U, s, V = _svd(Ft)
    e = np.log(np.maximum(np.abs(s), 1e-14))
    tr = e.sum()
    h = e - tr / 3
    hn = np.linalg.norm(h)
    dg = hn + (3 * lam + 2 * mu) / (2 * mu) * tr * alpha
    if dg <= 0:
        F = Ft
    elif tr > 0:
        F = U @ V.T
    else:
        F = U @ np.diag(np.exp(e - h * dg / hn)) @ V.T
    return F, _hencky(F, mu, lam)


def _von_mises(Ft, mu, lam, ys, xi=0.0):
    """(F, tau, yield stress after the map): hardening ys + 2 mu xi dg when yielding
    (mpm_utils.py:164-167)."""
    U, s, V = _svd(Ft)
    e = np.log(np.maximum(s, 0.01))
    t = 2 * mu * e + lam * e.sum()
    ys_next = ys
    if np.linalg.norm(t - t.mean()) > ys:
        h = e - e.mean()
        hn = np.linalg.norm(h) + 1e-6
        dg = hn - ys / (2 * mu)
        F = U @ np.diag(np.exp(e - dg / hn * h)) @ V.T
        ys_next = ys + 2 * mu * xi * dg
    else:
        F = Ft
    return F, _hencky(F, mu, lam), ys_next


def _mu_i(Ft, mu, lam, mu_s, dmu, I0, d, rho_s, dt):
    U, s, V = _svd(Ft)
    c = np.maximum(np.abs(s), 1e-6)
    e = np.log(c)
    tr = e.sum()
    if tr >= 0:
        F = U @ V.T
    else:
        me = tr / 3
        dv = e - me
        pK = -(2 * mu / 3 + lam) * tr
        tbK = SQ2 * mu * np.linalg.norm(dv)
        if tbK > mu_s * pK:
            Ic = d * np.sqrt(rho_s / (pK / c.prod()))
            lo, hi = 0.0, tbK / (mu * dt)
            for _ in range(200):
                mid = 0.5 * (lo + hi)
                I = mid * Ic
                if tbK - mu * dt * mid - (mu_s + dmu * I / (I + I0)) * pK > 0:
                    lo = mid
                else:
                    hi = mid
            k = (tbK - mu * dt * 0.5 * (lo + hi)) / tbK
            F = U @ np.diag(np.exp(dv * k + me)) @ V.T
        else:
            F = Ft
    return F, _hencky(F, mu, lam)


def _viscous(Ft, L, eta, bulk, ty, K, n):
    Jc = np.linalg.det(Ft) ** (1.0 / 3.0)
    J = Jc ** 3
    p = -bulk * (J ** -1.1 - 1.0)
    D = 0.5 * (L + L.T)
    D -= np.trace(D) / 3 * np.eye(3)
    gd = np.sqrt(2 * (D * D).sum() + 0.02 ** 2)
    t = J * (2 * (eta + ty / gd + K * gd ** (n - 1)) * D + p * np.eye(3))
    return Jc * np.eye(3), 0.5 * (t + t.T)


# ---------------------------------------------------------------- cases
_NAMES = ("mu", "lam", "ys", "eta", "bulk", "K", "n", "mu_s", "dmu", "I0", "d")
_PARAMS = {0: ("mu", "lam"), 1: ("mu", "lam", "ys"), 2: ("mu", "lam"), 9: ("mu", "lam", "mu_s", "dmu", "I0", "d"),
           10: ("eta", "bulk", "ys", "K", "n")}


def _rot(rng):
    q = np.linalg.qr(rng.standard_normal((3, 3)))[0]
    return q * np.sign(np.linalg.det(q))


def _cases(rng):
    out = []
    base = dict(dt=1e-4, mu=0.0, lam=0.0, ys=0.0, eta=0.0, bulk=0.0, K=0.0, n=1.0, mu_s=0.0, dmu=0.0, I0=1.0,
                d=0.0, rho_s=0.0)
    # fixed-corotated: E 1e5, nu 0.3; small and large stretches, nearly repeated singular
    # values, an inverted F (det F < 0)
    mu, lam = 1e5 / 2.6, 1e5 * 0.3 / (1.3 * 0.4)
    for kind in ("small", "large", "large", "repeated", "inverted"):
        if kind == "small":
            s = np.exp(rng.uniform(-0.02, 0.02, 3))
        elif kind == "large":
            s = np.exp(rng.uniform(-0.4, 0.4, 3))
        elif kind == "repeated":
            s = np.array([1.1, 1.1 + 1e-6, 0.9])
        else:
            s = np.array([1.2, 0.9, -0.3])
        out.append(dict(base, mat=0, mu=mu, lam=lam, Ft=_rot(rng) @ np.diag(s) @ _rot(rng).T))
    # Drucker-Prager sand: E 1e6, nu 0.3, friction angle 30 deg; inside the cone, on it
    # (compression), and in tension beyond it, each kept 2 % from the switch
    mu, lam = 1e6 / 2.6, 1e6 * 0.3 / (1.3 * 0.4)
    sphi = np.sin(np.radians(30.0))
    alpha = np.sqrt(2.0 / 3.0) * 2.0 * sphi / (3.0 - sphi)
    cc = (3 * lam + 2 * mu) / (2 * mu) * alpha
    for kind in ("elastic", "cone", "cone", "cone", "tension"):
        while True:
            tr = rng.uniform(0.002, 0.01) if kind == "tension" else -rng.uniform(0.005, 0.03)
            dv = rng.standard_normal(3)
            dv -= dv.mean()
            dv *= rng.uniform(0.001, 0.05) / np.linalg.norm(dv)
            r = np.linalg.norm(dv) / (-cc * tr) if tr < 0 else 2.0
            if kind == "tension" or (abs(r - 1) > 0.02 and (r > 1) == (kind == "cone")):
                break
        out.append(dict(base, mat=2, mu=mu, lam=lam, ys=alpha,
                        Ft=_rot(rng) @ np.diag(np.exp(tr / 3 + dv)) @ _rot(rng).T))
    # von Mises: E 2e5, nu 0.3, yield 1e3; elastic, plastic, near-repeated singular values,
    # with and without hardening (the eta and bulk slots carry xi and the adjoint of the
    # yield stress after the map, which every case passes through)
    mu, lam = 2e5 / 2.6, 2e5 * 0.3 / (1.3 * 0.4)
    for kind in ("elastic", "plastic", "plastic", "repeated", "repeated_plastic", "rotation",
                 "plastic_hard", "plastic_hard", "elastic_hard", "repeated_plastic_hard"):
        while True:
            if kind == "rotation":
                e = np.zeros(3)
            elif kind.startswith("repeated"):
                a = 0.0 if kind == "repeated" else 0.015
                # singular values 1e-6 apart around a common strain of 2-5e-3 (large
                # enough for float32 logs to resolve it to ~1e-5)
                e = np.array([a, -a + 1e-6, 2e-7]) + rng.choice([-1.0, 1.0]) * rng.uniform(2e-3, 5e-3)
            else:
                e = rng.uniform(-0.02, 0.02, 3)
            t = 2 * mu * e + lam * e.sum()
            r = np.linalg.norm(t - t.mean()) / 1e3
            if kind in ("rotation", "repeated") or (abs(r - 1) > 0.02 and (r > 1) == (not kind.startswith("elastic"))):
                break
        xi = rng.choice([0.5, 1.0]) if kind.endswith("_hard") else 0.0
        out.append(dict(base, mat=1, mu=mu, lam=lam, ys=1e3, eta=xi, bulk=rng.standard_normal() * 1e-3,
                        Ft=_rot(rng) @ np.diag(np.exp(e)) @ _rot(rng).T))
    # mu(I): E 1e6, nu 0.3; elastic / yielding compression and tension
    mu, lam = 1e6 / 2.6, 1e6 * 0.3 / (1.3 * 0.4)
    for kind in ("elastic", "yield", "yield", "yield", "tension"):
        while True:
            tr = rng.uniform(0.002, 0.01) if kind == "tension" else -rng.uniform(0.005, 0.03)
            dv = rng.standard_normal(3)
            dv -= dv.mean()
            dv *= rng.uniform(0.001, 0.05) / np.linalg.norm(dv)
            r = SQ2 * mu * np.linalg.norm(dv) / (0.38 * -(2 * mu / 3 + lam) * tr)
            if kind == "tension" or (abs(r - 1) > 0.02 and (r > 1) == (kind == "yield")):
                break
        out.append(dict(base, mat=9, mu=mu, lam=lam, mu_s=0.38, dmu=0.26, I0=0.28, d=1e-3, rho_s=2650.0,
                        Ft=_rot(rng) @ np.diag(np.exp(tr / 3 + dv)) @ _rot(rng).T))
    # Herschel-Bulkley fluid
    for _ in range(4):
        Ft = (1.0 + rng.uniform(-0.02, 0.02)) ** (1 / 3) * np.eye(3) + 0.01 * rng.uniform(-1, 1, (3, 3))
        out.append(dict(base, mat=10, eta=2.0, bulk=1e4, ys=50.0, K=20.0, n=0.5, Ft=Ft,
                        L=10.0 * rng.uniform(-1, 1, (3, 3))))
    for c in out:
        c.setdefault("L", np.zeros((3, 3)))
        c["Fb"] = rng.standard_normal((3, 3))
        c["tb"] = rng.standard_normal((3, 3)) / max(c["mu"], c["bulk"] if c["mat"] != 1 else 0.0, 1.0)
        for k in ("Ft", "L", "Fb", "tb"):  # the kernel sees these in float32
            c[k] = c[k].astype(np.float32).astype(np.float64)
        for k in _NAMES + ("dt", "rho_s"):
            c[k] = float(np.float32(c[k]))
    return out


def _objective(c, stress_only=False, **over):
    p = dict(c, **over)
    if p["mat"] == 0:
        F, t = _fcr(p["Ft"], p["mu"], p["lam"])
    elif p["mat"] == 2:
        F, t = _sand(p["Ft"], p["mu"], p["lam"], p["ys"])
    elif p["mat"] == 1:
        F, t, ys_next = _von_mises(p["Ft"], p["mu"], p["lam"], p["ys"], p["eta"])
        if not stress_only:  # + (adjoint of the yield stress after the map) * that stress
            return (F * p["Fb"]).sum() + (t * p["tb"]).sum() + p["bulk"] * ys_next
    elif p["mat"] == 9:
        F, t = _mu_i(p["Ft"], p["mu"], p["lam"], p["mu_s"], p["dmu"], p["I0"], p["d"], p["rho_s"], p["dt"])
    else:
        F, t = _viscous(p["Ft"], p["L"], p["eta"], p["bulk"], p["ys"], p["K"], p["n"])
    return (t * p["tb"]).sum() if stress_only else (F * p["Fb"]).sum() + (t * p["tb"]).sum()


def _fd_matrix(c, key, h):
    G = np.zeros((3, 3))
    for i in range(3):
        for j in range(3):
            E = np.zeros((3, 3))
            E[i, j] = h
            G[i, j] = (_objective(c, **{key: c[key] + E}) - _objective(c, **{key: c[key] - E})) / (2 * h)
    return G


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc not found")
    exe = str(tmp_path_factory.mktemp("adjoint_math") / "adjoint_math_check")
    cmd = [HIPCC, "--offload-arch=gfx950", "-O2", "-std=c++17", "-Xclang", "-target-feature", "-Xclang",
           "-packed-fp32-ops", "-I", os.path.join(ROOT, "include"), "-I", os.path.join(ROOT, "mpm-engine_amd", "csrc"),
           os.path.join(ROOT, "tests", "host", "adjoint_math_check.hip"), "-o", exe]
    subprocess.run(cmd, check=True, capture_output=True)
    return exe


def test_constitutive_adjoints_match_float64_differences(driver):
    rng = np.random.default_rng(17)
    cases = _cases(rng)
    lines = []
    for c in cases:
        # input order: mat dt mu lam ys eta bulk K n mu_s dmu I0 d rho_s
        head = [c["mat"], c["dt"], c["mu"], c["lam"], c["ys"], c["eta"], c["bulk"], c["K"], c["n"], c["mu_s"],
                c["dmu"], c["I0"], c["d"], c["rho_s"]]
        vals = head + list(c["Ft"].ravel()) + list(c["L"].ravel()) + list(c["Fb"].ravel()) + list(c["tb"].ravel())
        lines.append(" ".join(f"{v:.9g}" for v in vals))
    res = subprocess.run([driver], input="\n".join(lines) + "\n", capture_output=True, text=True, check=True)
    rows = [np.array(r.split(), float) for r in res.stdout.strip().splitlines()]
    assert len(rows) == len(cases)
    order = ("mu", "lam", "ys", "eta", "bulk", "K", "n", "mu_s", "dmu", "I0", "d")
    for idx, (c, row) in enumerate(zip(cases, rows)):
        got = dict(zip(order, row[18:]))
        fd_F = _fd_matrix(c, "Ft", 1e-7)
        err = np.abs(row[:9].reshape(3, 3) - fd_F).max() / np.abs(fd_F).max()
        assert err < 2e-4, f"case {idx} (mat {c['mat']}): F_trial adjoint off by {err:.2e}"
        if c["mat"] == 10:
            fd_L = _fd_matrix(c, "L", 1e-6)
            err = np.abs(row[9:18].reshape(3, 3) - fd_L).max() / np.abs(fd_L).max()
            assert err < 2e-4, f"case {idx}: L adjoint off by {err:.2e}"
        # parameter derivatives are held against the size of the stress term they scale,
        # |<tau, tb>| / |theta|: where the derivative itself is a cancelling sum (nearly
        # equal singular values) float32 rounding of its terms is what remains, while a
        # wrong or dropped term is off by O(1) of that size
        # (plus the stress a strain of 1e-6, the resolution of float32 logs near identity,
        # makes against tb: the rotation case has exactly zero derivatives)
        jt = abs(_objective(c, stress_only=True))
        floor = 1e-6 * max(c["mu"], c["lam"], c["bulk"], 1.0) * np.abs(c["tb"]).sum()
        for name in _PARAMS[c["mat"]]:
            h = 1e-6 * abs(c[name])
            fd = (_objective(c, **{name: c[name] + h}) - _objective(c, **{name: c[name] - h})) / (2 * h)
            tol = 1e-3 * max(abs(fd), jt / abs(c[name])) + floor / abs(c[name])
            assert abs(got[name] - fd) <= tol, \
                f"case {idx} (mat {c['mat']}): d/d {name} {got[name]:.6e} vs {fd:.6e}"
