"""Coupling helpers on the CPU: the stress-integral box wrench estimator against closed
forms (coupling/wrench.py:15-46 of the reference), the force admittance law
(coupling/admittance.py:14-43), and the dough + sand press scene's composition."""
from __future__ import annotations

import numpy as np

from mpm_amd.coupling import ForceAdmittance, box_contact_wrench


def _layer(p=1000.0, n_side=10, T=0.01, offset=(0.0, 0.0)):
    """n_side^2 x 3 particles filling the contact layer under a box at z = 0.5 bottom."""
    g = np.stack(np.meshgrid(np.linspace(-0.04, 0.04, n_side), np.linspace(-0.04, 0.04, n_side),
                             np.linspace(0.5 - 0.9 * T, 0.5 + 0.4 * T, 3), indexing="ij"), -1).reshape(-1, 3)
    g[:, 0] += offset[0]
    g[:, 1] += offset[1]
    sig = np.tile(-p * np.eye(3), (len(g), 1, 1))
    return g, sig


def test_hydrostatic_layer_closed_form():
    x, sig = _layer()
    vol = np.full(len(x), 2e-6)
    T = 0.01
    w = box_contact_wrench(x, sig, vol, (0.0, 0.0, 0.55), (0.05, 0.05, 0.05), layer_cells=1.0, dx=T)
    assert w["n_contact"] == len(x)
    assert np.isclose(w["Fz"], 1000.0 * len(x) * 2e-6 / T)   # -(sigma . e_z) vol / T, upward
    assert np.allclose(w["force"][:2], 0.0) and np.allclose(w["torque"], 0.0, atol=1e-12)


def test_offset_layer_torque_and_tension_gate():
    x, sig = _layer(offset=(0.005, 0.0))
    vol = np.full(len(x), 2e-6)
    w = box_contact_wrench(x, sig, vol, (0.0, 0.0, 0.55), (0.05, 0.05, 0.05), layer_cells=1.0, dx=0.01)
    r = x - np.array([0.0, 0.0, 0.55])
    f = np.tile([0.0, 0.0, 1000.0 * 2e-6 / 0.01], (len(x), 1))
    assert np.allclose(w["torque"], np.cross(r, f).sum(0))
    w2 = box_contact_wrench(x, -sig, vol, (0.0, 0.0, 0.55), (0.05, 0.05, 0.05), layer_cells=1.0, dx=0.01)
    assert w2["n_contact"] == 0 and w2["Fz"] == 0.0      # tension only: gated out
    w3 = box_contact_wrench(x, -sig, vol, (0.0, 0.0, 0.55), (0.05, 0.05, 0.05), layer_cells=1.0, dx=0.01,
                            compressive_only=False)
    assert np.isclose(w3["Fz"], -w["Fz"])


def test_default_layer_is_three_particle_sizes():
    x, sig = _layer(T=0.006)
    vol = np.full(len(x), 8e-9)  # h = 2 mm -> T = 6 mm
    w = box_contact_wrench(x, sig, vol, (0.0, 0.0, 0.55), (0.05, 0.05, 0.05))
    assert np.isclose(w["Fz"], 1000.0 * len(x) * 8e-9 / 0.006)


def test_force_admittance():
    a = ForceAdmittance(f_target=30.0, v_max=0.05)
    assert a.velocity_down(0.0) == 0.05 and a.velocity_down(30.0) == 0.0
    assert np.isclose(a.velocity_down(15.0), 0.025) and a.velocity_down(90.0) == -0.05
    assert ForceAdmittance(30.0, 0.05, allow_retract=False).velocity_down(90.0) == 0.0


def test_dough_sand_press_composition():
    from mpm_amd.scenes import DoughSandPress

    sc = DoughSandPress()
    n_sand = len(sc.pos) - sc.n_dough
    assert 290_000 < sc.n_dough < 310_000 and 290_000 < n_sand < 310_000
    assert sc.pos[:, 2].min() > sc.floor and np.isfinite(sc.pos).all()
    d, s = sc.pos[:sc.n_dough], sc.pos[sc.n_dough:]
    assert d[:, 0].max() < s[:, 0].min()          # slab and heap side by side
