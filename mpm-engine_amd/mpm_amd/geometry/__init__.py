"""Signed-distance-field containers for the mesh-SDF and CDF (thin-boundary) colliders.

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


# ------------------------------------------------------------------ CDF (open surfaces)
@dataclass
class CDFData:
    """Colored distance field of an OPEN oriented mid-surface for the CPIC thin-boundary
    collider (the reference's geometry.CDFData, src/warpmpm/geometry/mesh_sdf.py:48-73):
    values[res, res, res] = distance to the sheet signed by the side of the nearest
    face's normal, valid[res, res, res] = 1 where distance and side are trustworthy
    (within `band` of the sheet and farther than `band` from its open rim), origin (3,),
    cell (isotropic voxel size) and band, in the body frame, z fastest."""

    values: np.ndarray
    valid: np.ndarray
    origin: np.ndarray
    cell: float
    band: float

    @property
    def res(self) -> int:
        return int(self.values.shape[0])


def _triangle_samples(verts, faces, spacing):
    """Points on every triangle no farther apart than about `spacing`: a lattice along the
    shortest edge and the other edge leaving one of its ends (steps sized per edge, so
    long thin triangles stay cheap) plus the third edge, with the index of the face each
    point lies on."""
    pts, owner = [], []
    for fi, tri in enumerate(verts[faces]):
        e = [np.linalg.norm(tri[(k + 1) % 3] - tri[k]) for k in range(3)]
        k = int(np.argmin(e))  # shortest edge tri[k] -> tri[k+1]
        a, b, c = tri[k], tri[(k + 1) % 3], tri[(k + 2) % 3]
        mu = max(1, int(np.ceil(np.linalg.norm(b - a) / spacing)))
        mv = max(1, int(np.ceil(np.linalg.norm(c - a) / spacing)))
        i, j = np.meshgrid(np.arange(mu + 1), np.arange(mv + 1), indexing="ij")
        u, v = (i / mu).ravel(), (j / mv).ravel()
        keep = u + v <= 1.0 + 1e-12
        u, v = u[keep], v[keep]
        mw = max(1, int(np.ceil(np.linalg.norm(c - b) / spacing)))
        t = np.arange(mw + 1) / mw
        u = np.concatenate([u, 1.0 - t])
        v = np.concatenate([v, t])
        pts.append(a[None] * (1.0 - u - v)[:, None] + b[None] * u[:, None] + c[None] * v[:, None])
        owner.append(np.full(len(u), fi, dtype=np.int64))
    return np.concatenate(pts), np.concatenate(owner)


def open_edges(faces) -> np.ndarray:
    """Edges used by exactly one face (the open rim of the surface), (E, 2) vertex ids."""
    e = np.sort(np.concatenate([faces[:, [0, 1]], faces[:, [1, 2]], faces[:, [2, 0]]]), axis=1)
    uniq, count = np.unique(e, axis=0, return_counts=True)
    return uniq[count == 1]


def _edge_samples(verts, edges, spacing):
    if len(edges) == 0:
        return np.zeros((0, 3))
    out = []
    for a, b in verts[edges]:
        m = max(1, int(np.ceil(np.linalg.norm(b - a) / spacing)))
        t = np.arange(m + 1) / m
        out.append(a[None] * (1.0 - t)[:, None] + b[None] * t[:, None])
    return np.concatenate(out)


def build_surface_cdf(verts, faces, res: int = 64, band_cells: float = 3.0,
                      margin_cells: float = 4.0) -> CDFData:
    """Voxelise an open oriented triangle surface into a CDFData.

    The cubic voxel grid spans the surface's bounding box plus `margin_cells` cells each
    side. Each voxel takes the distance to the nearest point of a dense sample of the
    surface (samples no farther apart than 0.4 cells) and the sign of its offset along
    that point's face normal; it is valid when that distance is within band =
    band_cells * cell and the distance to the open rim (edges used by one face) exceeds
    the band, since no consistent side exists past the rim. Voxels farther than
    2 band + 2 cells from the surface are invalid and hold that distance, positive."""
    from scipy.spatial import cKDTree

    verts = np.asarray(verts, dtype=np.float64)
    faces = np.asarray(faces, dtype=np.int64)
    lo, hi = verts.min(axis=0), verts.max(axis=0)
    cell = float((hi - lo).max()) / (res - 1 - 2 * margin_cells)
    band = band_cells * cell
    origin = 0.5 * (lo + hi) - 0.5 * cell * (res - 1)
    ax = [origin[d] + cell * np.arange(res) for d in range(3)]
    grid = np.stack(np.meshgrid(*ax, indexing="ij"), axis=-1).reshape(-1, 3)
    fa, fb, fc = verts[faces[:, 0]], verts[faces[:, 1]], verts[faces[:, 2]]
    normals = np.cross(fb - fa, fc - fa)
    normals /= np.maximum(np.linalg.norm(normals, axis=1, keepdims=True), 1e-30)
    samples, owner = _triangle_samples(verts, faces, 0.4 * cell)
    # distances past `reach` only ever read as invalid, so the nearest-sample queries
    # stop there (far voxels get +reach)
    reach = 2.0 * band + 2.0 * cell
    dist, idx = cKDTree(samples).query(grid, distance_upper_bound=reach, workers=-1)
    near = np.isfinite(dist)
    rim = _edge_samples(verts, open_edges(faces), 0.4 * cell)
    rim_dist = (cKDTree(rim).query(grid, distance_upper_bound=reach, workers=-1)[0] if len(rim)
                else np.full(len(grid), np.inf))
    side = np.ones(len(grid))
    j = idx[near]
    side[near] = np.sign(np.einsum("ij,ij->i", grid[near] - samples[j], normals[owner[j]]))
    side[side == 0.0] = 1.0
    dist = np.where(near, dist, reach)
    values = (side * dist).reshape(res, res, res)
    valid = ((dist <= band) & (rim_dist > band)).astype(np.float64).reshape(res, res, res)
    return CDFData(values=values, valid=valid, origin=origin, cell=cell, band=band)


def square_sheet(size: float, normal_axis: int = 2):
    """Two-triangle square sheet of side `size` through the origin, normal along
    +axis (0 = x, 2 = z): (verts, faces)."""
    s = 0.5 * size
    if normal_axis == 0:
        v = np.array([[0, -s, -s], [0, s, -s], [0, s, s], [0, -s, s]], dtype=float)
    elif normal_axis == 1:
        v = np.array([[-s, 0, -s], [-s, 0, s], [s, 0, s], [s, 0, -s]], dtype=float)
    else:
        v = np.array([[-s, -s, 0], [s, -s, 0], [s, s, 0], [-s, s, 0]], dtype=float)
    return v, np.array([[0, 1, 2], [0, 2, 3]], dtype=np.int64)


def revolve_profile_open(profile, n_theta: int = 64):
    """Surface of revolution of an open (r, z) polyline about the z axis, normals
    pointing away from the axis for a profile walked upward: (verts, faces)."""
    prof = np.asarray(profile, dtype=np.float64)
    th = 2.0 * np.pi * np.arange(n_theta) / n_theta
    verts = np.array([[r * np.cos(t), r * np.sin(t), z] for r, z in prof for t in th])
    faces = []
    for i in range(len(prof) - 1):
        for j in range(n_theta):
            a, b = i * n_theta + j, i * n_theta + (j + 1) % n_theta
            c, d = a + n_theta, b + n_theta
            faces += [[a, b, d], [a, d, c]]
    return verts, np.asarray(faces, dtype=np.int64)


def save_cdf(cdf: CDFData, path) -> None:
    np.savez_compressed(path, values=cdf.values, valid=cdf.valid, origin=cdf.origin,
                        cell=cdf.cell, band=cdf.band)


def load_cdf(path) -> CDFData:
    d = np.load(path)
    return CDFData(values=d["values"], valid=d["valid"], origin=d["origin"],
                   cell=float(d["cell"]), band=float(d["band"]))


__all__ = ["SDFData", "sdf_from_function", "sphere_sdf", "CDFData", "build_surface_cdf",
           "square_sheet", "revolve_profile_open", "open_edges", "save_cdf", "load_cdf"]
