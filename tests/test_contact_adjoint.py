"""The contact maps of the revolved-cup and mesh-SDF colliders and their transposed
Jacobians (csrc/mpm_adjoint_math.h: surface_apply, surface_apply_T), used by the grid
adjoint, against a float64 restatement of the reference's rules and its central
differences, on the CPU.

The velocity map at a node in contact, for the surface velocity vs and unit normal n:
grab (v = vs, the cup's sticky layer and the mesh collider's sticky surface), slip (the
normal part of v - vs removed), separable Coulomb on the relative velocity (v unchanged
when separating, else vs plus the tangential part shortened by friction * |normal part|
and stopped when that exceeds it) in the cup's and the mesh collider's expressions
(mpm_solver_warp.py:2161-2225 and :2680-2736), and a cup node without a normal (passes).
Cases keep 1e-3 away from the switches (separation, stopping), where the map has a kink.
The GPU tests check the kernels end to end, composed with the impulse and torque terms."""
from __future__ import annotations

import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"


def _contact(mode, cup, vs, n, fr, vin):
    """float64 restatement of the contact rules (mpm_solver_warp.py:2161-2225, 2680-2736)."""
    if mode == 1:
        return vs.copy()
    if mode == 4:
        return vin.copy()
    vr = vin - vs
    vn = vr @ n
    if mode == 3:
        return vs + (vr - n * vn)
    if not vn < 0.0:
        return vin.copy()
    vt = vr - n * vn
    tl = np.linalg.norm(vt)
    if cup:
        if tl > 1e-20:
            vt = vt / tl * max(tl + fr * vn, 0.0)
    elif tl > 1e-12:
        vt = vt * (max(0.0, tl + fr * vn) / tl)
    return vs + vt


def _cases(rng):
    out = []
    for cup in (1, 0):
        for kind in ("grab", "pass", "slip", "separating", "sliding", "sliding", "stopped", "frictionless"):
            if kind == "pass" and not cup:
                continue
            while True:
                n = rng.standard_normal(3)
                n /= np.linalg.norm(n)
                vs = rng.uniform(-0.3, 0.3, 3)
                vin = vs + rng.uniform(-1.0, 1.0, 3)
                fr = {"frictionless": 0.0}.get(kind, rng.choice([0.05, 0.4, 1.5]))
                vr = vin - vs
                vn = vr @ n
                tl = np.linalg.norm(vr - n * vn)
                mode = {"grab": 1, "pass": 4, "slip": 3}.get(kind, 2)
                if kind == "separating" and vn > 1e-3:
                    break
                if kind in ("sliding", "frictionless") and vn < -1e-3 and tl + fr * vn > 1e-3:
                    break
                if kind == "stopped" and vn < -1e-3 and tl + fr * vn < -1e-3:
                    break
                if kind in ("grab", "pass", "slip"):
                    break
            f32 = lambda a: np.asarray(a, np.float32).astype(np.float64)  # noqa: E731
            out.append(dict(mode=mode, cup=cup, vs=f32(vs), n=f32(n), fr=float(np.float32(fr)), vin=f32(vin),
                            g=f32(rng.standard_normal(3)), kind=kind))
    return out


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc not found")
    exe = str(tmp_path_factory.mktemp("contact_adjoint") / "contact_adjoint_check")
    cmd = [HIPCC, "--offload-arch=gfx950", "-O2", "-std=c++17", "-Xclang", "-target-feature", "-Xclang",
           "-packed-fp32-ops", "-I", os.path.join(ROOT, "include"), "-I", os.path.join(ROOT, "mpm-engine_amd", "csrc"),
           os.path.join(ROOT, "tests", "host", "contact_adjoint_check.hip"), "-o", exe]
    subprocess.run(cmd, check=True, capture_output=True)
    return exe


def test_contact_maps_and_transposed_jacobians(driver):
    rng = np.random.default_rng(31)
    cases = [c for _ in range(6) for c in _cases(rng)]
    lines = [" ".join(f"{v:.9g}" for v in [c["mode"], c["cup"], *c["vs"], *c["n"], c["fr"], *c["vin"], *c["g"]])
             for c in cases]
    res = subprocess.run([driver], input="\n".join(lines) + "\n", capture_output=True, text=True, check=True)
    rows = [np.array(r.split(), float) for r in res.stdout.strip().splitlines()]
    assert len(rows) == len(cases)
    for idx, (c, row) in enumerate(zip(cases, rows)):
        want = _contact(c["mode"], c["cup"], c["vs"], c["n"], c["fr"], c["vin"])
        scale = np.abs(c["vin"]).max() + np.abs(c["vs"]).max()
        assert np.abs(row[:3] - want).max() <= 1e-6 * scale, f"case {idx} ({c['kind']}): forward {row[:3]} vs {want}"
        # J^T g against central differences of the float64 map
        h = 1e-7
        J = np.zeros((3, 3))
        for k in range(3):
            e = np.zeros(3)
            e[k] = h
            J[:, k] = (_contact(c["mode"], c["cup"], c["vs"], c["n"], c["fr"], c["vin"] + e)
                       - _contact(c["mode"], c["cup"], c["vs"], c["n"], c["fr"], c["vin"] - e)) / (2 * h)
        jt = J.T @ c["g"]
        tol = 2e-5 * (np.abs(jt).max() + np.abs(c["g"]).max())
        assert np.abs(row[3:] - jt).max() <= tol, f"case {idx} ({c['kind']}, cup {c['cup']}): {row[3:]} vs {jt}"
    kinds = {c["kind"] for c in cases}
    assert {"grab", "pass", "slip", "separating", "sliding", "stopped", "frictionless"} <= kinds
