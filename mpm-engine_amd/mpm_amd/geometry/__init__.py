"""Signed-distance-field containers for the mesh-SDF collider.

``SDFData`` has the reference's field layout (src/warpmpm/geometry/mesh_sdf.py:26-45):
values[res, res, res] (negative inside), grads[res, res, res, 3], origin (3,), cell, sdf_max,
all in the collider body frame, z fastest. The mesh -> SDF winding-number builder of the
reference is not part of this package; ``sdf_from_function`` samples any analytic SDF
(and its central-difference gradient) onto that layout.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass
class SDFData:
    values: np.ndarray
    grads: np.ndarray
    origin: np.ndarray
    cell: float
    sdf_max: float

    @property
    def res(self) -> int:
        return int(self.values.shape[0])


def sdf_from_function(fn, lo, hi, res: int) -> SDFData:
    """Sample fn(points[..., 3]) -> signed distance on a cubic res^3 voxel grid spanning
    [lo, lo + (res - 1) * cell] with cell = max(hi - lo) / (res - 1)."""
    lo = np.asarray(lo, dtype=np.float64)
    hi = np.asarray(hi, dtype=np.float64)
    cell = float((hi - lo).max()) / (res - 1)
    ax = np.arange(res) * cell
    pts = np.stack(np.meshgrid(lo[0] + ax, lo[1] + ax, lo[2] + ax, indexing="ij"), axis=-1)
    vals = fn(pts)
    h = 0.5 * cell
    grads = np.stack([(fn(pts + e) - fn(pts - e)) / (2.0 * h)
                      for e in (np.array([h, 0, 0]), np.array([0, h, 0]), np.array([0, 0, h]))],
                     axis=-1)
    return SDFData(values=vals.astype(np.float32), grads=grads.astype(np.float32),
                   origin=lo.astype(np.float32), cell=cell, sdf_max=float(np.abs(vals).max()))


def sphere_sdf(radius: float, res: int = 32, margin: float = 4.0) -> SDFData:
    """Solid ball of `radius` about the body origin, with `margin` cells around it."""
    ext = radius * (1.0 + 2.0 * margin / (res - 1))
    fn = lambda p: np.linalg.norm(p, axis=-1) - radius  # noqa: E731
    return sdf_from_function(fn, [-ext] * 3, [ext] * 3, res)


__all__ = ["SDFData", "sdf_from_function", "sphere_sdf"]
