"""Stress-integral estimate of the reaction wrench on a box end-effector from exported
particle arrays (src/warpmpm/coupling/wrench.py:15-46): the traction -(sigma . e_z) of the
particles in a contact layer of thickness T under the box bottom, integrated as
volume / T. The same estimator runs on the device through `Solver.box_contact_wrench`
(mpm_box_contact_wrench), which avoids exporting the particle state."""
from __future__ import annotations

import numpy as np


def box_contact_wrench(x, cauchy, vol, box_center, box_half, layer_cells: float | None = None,
                       dx: float | None = None, compressive_only: bool = True) -> dict:
    """x (N, 3), cauchy (N, 3, 3), vol (N,): current particle positions, Cauchy stress and
    volume. T = layer_cells * dx when both are given, else 3 mean(vol)^(1/3). Returns
    {'force', 'torque' (about the box centre), 'n_contact', 'Fz'}; zeros below 3 contacts.
    With compressive_only only particles with sigma_zz < 0 count."""
    x = np.asarray(x, dtype=np.float64)
    sig = np.asarray(cauchy, dtype=np.float64)
    vol = np.asarray(vol, dtype=np.float64)
    c = np.asarray(box_center, dtype=np.float64)
    half = np.asarray(box_half, dtype=np.float64)
    T = float(layer_cells) * float(dx) if (layer_cells and dx) else 3.0 * float(np.mean(vol)) ** (1.0 / 3.0)
    bottom = c[2] - half[2]
    sel = ((np.abs(x[:, 0] - c[0]) < half[0]) & (np.abs(x[:, 1] - c[1]) < half[1])
           & (x[:, 2] > bottom - T) & (x[:, 2] < bottom + 0.5 * T))
    if compressive_only:
        sel &= sig[:, 2, 2] < 0.0
    n = int(np.count_nonzero(sel))
    if n < 3:
        return {"force": np.zeros(3), "torque": np.zeros(3), "n_contact": 0, "Fz": 0.0}
    f = -sig[sel][:, :, 2] * (vol[sel] / T)[:, None]
    force = f.sum(axis=0)
    torque = np.cross(x[sel] - c, f).sum(axis=0)
    return {"force": force, "torque": torque, "n_contact": n, "Fz": float(force[2])}
