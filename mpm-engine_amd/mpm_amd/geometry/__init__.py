"""Signed-distance-field containers for the mesh-SDF and CDF (thin-boundary) colliders.

``SDFData`` has the reference's field layout (src/warpmpm/geometry/mesh_sdf.py:26-45):
values[res, res, res] (negative inside), grads[res, res, res, 3], origin (3,), cell, sdf_max,
all in the collider body frame, z fastest. ``build_sdf`` voxelises a closed triangle mesh
(unsigned distance + winding-number sign, the reference's builder,
src/warpmpm/geometry/mesh_sdf.py:328-366), ``build_surface_cdf`` an open sheet
(:412-481); both run on the host (numpy, device="cpu") or in the HIP kernels
mpm_mesh_field / mpm_segment_distance (device="hip", exact distances).
``sdf_from_function`` samples any analytic SDF onto the same layout.
"""
from __future__ import annotations

import hashlib
from dataclasses import dataclass
from pathlib import Path

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


def _cubic_grid(verts, res, margin_cells):
    lo, hi = verts.min(axis=0), verts.max(axis=0)
    cell = float((hi - lo).max()) / (res - 1 - 2 * margin_cells)
    origin = 0.5 * (lo + hi) - 0.5 * cell * (res - 1)
    ax = [origin[d] + cell * np.arange(res) for d in range(3)]
    return origin, cell, np.stack(np.meshgrid(*ax, indexing="ij"), axis=-1).reshape(-1, 3)


def winding_number(points, verts, faces, chunk: int = 4096) -> np.ndarray:
    """Generalised winding number of a triangle mesh at each point: the sum of the solid
    angles of its triangles (Van Oosterom & Strackee 1983) over 4 pi; ~1 inside a closed
    outward-oriented mesh, ~0 outside."""
    pts = np.asarray(points, dtype=np.float64).reshape(-1, 3)
    tri = np.asarray(verts, dtype=np.float64)[np.asarray(faces, dtype=np.int64)]
    out = np.empty(len(pts))
    for s in range(0, len(pts), chunk):
        p = pts[s:s + chunk, None, None, :]
        x, y, z = tri[None, :, 0] - p[:, :, 0], tri[None, :, 1] - p[:, :, 0], tri[None, :, 2] - p[:, :, 0]
        lx, ly, lz = (np.linalg.norm(v, axis=-1) for v in (x, y, z))
        num = np.einsum("pfi,pfi->pf", x, np.cross(y, z))
        den = (lx * ly * lz + np.einsum("pfi,pfi->pf", x, y) * lz + np.einsum("pfi,pfi->pf", x, z) * ly
               + np.einsum("pfi,pfi->pf", y, z) * lx)
        out[s:s + chunk] = (2.0 * np.arctan2(num, den)).sum(axis=1) / (4.0 * np.pi)
    return out


def _hip_field(verts, faces, origin, cell, res, mode):
    """Run mpm_mesh_field on the GPU; returns (dist, aux) as float64 res^3 arrays."""
    import ctypes as C

    import torch

    from ..core import native

    lib = native.load()
    dev = torch.device("cuda")
    v = torch.from_numpy(np.ascontiguousarray(verts, dtype=np.float32).reshape(-1)).to(dev)
    f = torch.from_numpy(np.ascontiguousarray(faces, dtype=np.int32).reshape(-1)).to(dev)
    dist = torch.empty(res ** 3, dtype=torch.float32, device=dev)
    aux = torch.empty(res ** 3, dtype=torch.float32, device=dev)
    o = (C.c_float * 3)(*[float(t) for t in origin])
    stream = torch.cuda.current_stream(dev).cuda_stream
    native.check(lib.mpm_mesh_field(native.ptr(v), len(verts), native.ptr(f), len(faces), o, float(cell), int(res),
                                    int(mode), native.ptr(dist), native.ptr(aux), stream), "mesh_field")
    return (dist.cpu().numpy().astype(np.float64).reshape(res, res, res),
            aux.cpu().numpy().astype(np.float64).reshape(res, res, res))


def _hip_segments(segments, origin, cell, res):
    import ctypes as C

    import torch

    from ..core import native

    lib = native.load()
    dev = torch.device("cuda")
    sg = torch.from_numpy(np.ascontiguousarray(segments, dtype=np.float32).reshape(-1)).to(dev)
    dist = torch.empty(res ** 3, dtype=torch.float32, device=dev)
    o = (C.c_float * 3)(*[float(t) for t in origin])
    stream = torch.cuda.current_stream(dev).cuda_stream
    native.check(lib.mpm_segment_distance(native.ptr(sg), len(segments), o, float(cell), int(res), native.ptr(dist),
                                          stream), "segment_distance")
    return dist.cpu().numpy().astype(np.float64).reshape(res, res, res)


def build_sdf(verts, faces, res: int = 64, margin_cells: float = 4.0, interior_probe=None,
              device: str = "cpu") -> SDFData:
    """Voxelise a closed triangle mesh into an SDFData (negative inside) on a cubic grid
    spanning its bounding box plus `margin_cells` cells (mesh_sdf.py:328-366): unsigned
    distance to the surface, sign from the winding number, the global orientation fixed by
    the winding number at `interior_probe` (default the bounding-box centre). device="cpu":
    distances to a dense surface sample (0.4-cell spacing), numpy winding numbers;
    device="hip": exact point-triangle distances and winding numbers in mpm_mesh_field."""
    verts = np.asarray(verts, dtype=np.float64)
    faces = np.asarray(faces, dtype=np.int64)
    if faces.size == 0 or faces.min() < 0 or faces.max() >= len(verts):
        raise ValueError("faces must index the vertex array")
    origin, cell, grid = _cubic_grid(verts, res, margin_cells)
    if device == "hip":
        dist, wn = _hip_field(verts, faces, origin, cell, res, mode=0)
    elif device == "cpu":
        from scipy.spatial import cKDTree

        samples, _ = _triangle_samples(verts, faces, 0.4 * cell)
        dist = cKDTree(samples).query(grid, workers=-1)[0].reshape(res, res, res)
        wn = winding_number(grid, verts, faces).reshape(res, res, res)
    else:
        raise ValueError("device must be 'cpu' or 'hip'")
    lo, hi = verts.min(axis=0), verts.max(axis=0)
    probe = 0.5 * (lo + hi) if interior_probe is None else np.asarray(interior_probe, dtype=np.float64)
    w_probe = float(winding_number(probe[None], verts, faces)[0])
    inside = (wn * np.sign(w_probe)) > 0.5 if abs(w_probe) > 0.1 else np.abs(wn) > 0.5
    signed = np.where(inside, -dist, dist)
    grads = np.stack(np.gradient(signed, cell, edge_order=2), axis=-1)
    return SDFData(values=signed, grads=grads, origin=origin, cell=float(cell),
                   sdf_max=float(np.abs(signed).max()))


def icosphere(radius: float = 1.0, subdivisions: int = 2):
    """Triangulated sphere (outward faces): (verts, faces)."""
    t = (1.0 + 5.0 ** 0.5) / 2.0
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
         (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11),
         (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    verts = [np.array(p, dtype=np.float64) / np.linalg.norm(p) for p in v]
    faces = list(f)
    for _ in range(subdivisions):
        cache, nf = {}, []

        def mid(a, b):
            key = (min(a, b), max(a, b))
            if key not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[key] = len(verts) - 1
            return cache[key]

        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    return radius * np.array(verts), np.array(faces, dtype=np.int64)


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
                      margin_cells: float = 4.0, device: str = "cpu") -> CDFData:
    """Voxelise an open oriented triangle surface into a CDFData.

    The cubic voxel grid spans the surface's bounding box plus `margin_cells` cells each
    side. Each voxel takes the distance to the nearest point of a dense sample of the
    surface (samples no farther apart than 0.4 cells) and the sign of its offset along
    that point's face normal; it is valid when that distance is within band =
    band_cells * cell and the distance to the open rim (edges used by one face) exceeds
    the band, since no consistent side exists past the rim. Voxels farther than
    2 band + 2 cells from the surface are invalid and hold that distance, positive.
    device="hip" computes exact point-triangle distances, the nearest triangle's side and
    the rim distance in mpm_mesh_field / mpm_segment_distance instead (no reach cut-off)."""
    from scipy.spatial import cKDTree

    verts = np.asarray(verts, dtype=np.float64)
    faces = np.asarray(faces, dtype=np.int64)
    if device == "hip":
        if faces.size == 0 or faces.min() < 0 or faces.max() >= len(verts):
            raise ValueError("faces must index the vertex array")
        origin, cell, _ = _cubic_grid(verts, res, margin_cells)
        band = band_cells * cell
        dist, side = _hip_field(verts, faces, origin, cell, res, mode=1)
        edges = open_edges(faces)
        rim = (_hip_segments(verts[edges].reshape(-1, 6), origin, cell, res) if len(edges)
               else np.full((res, res, res), np.inf))
        valid = ((dist <= band) & (rim > band)).astype(np.float64)
        return CDFData(values=side * dist, valid=valid, origin=origin, cell=cell, band=band)
    if device != "cpu":
        raise ValueError("device must be 'cpu' or 'hip'")
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
    path = Path(path)
    path.parent.mkdir(parents=True, exist_ok=True)
    np.savez_compressed(path, values=cdf.values, valid=cdf.valid, origin=cdf.origin,
                        cell=cdf.cell, band=cdf.band)


def load_cdf(path) -> CDFData:
    d = np.load(path)
    return CDFData(values=d["values"], valid=d["valid"], origin=d["origin"],
                   cell=float(d["cell"]), band=float(d["band"]))


def save_sdf(sdf: SDFData, path) -> None:
    """npz of an SDFData (mesh_sdf.py:538-542 layout: values, grads, origin, cell, sdf_max)."""
    path = Path(path)
    path.parent.mkdir(parents=True, exist_ok=True)
    np.savez_compressed(path, values=sdf.values, grads=sdf.grads, origin=sdf.origin,
                        cell=sdf.cell, sdf_max=sdf.sdf_max)


def load_sdf(path) -> SDFData:
    d = np.load(path)
    return SDFData(values=d["values"], grads=d["grads"], origin=d["origin"],
                   cell=float(d["cell"]), sdf_max=float(d["sdf_max"]))


# Versions of the cached field formats; bump when a builder's output changes.
_SDF_CACHE_VERSION = 1
_CDF_CACHE_VERSION = 1


def _field_key(version, verts, faces, params, device, probe=None) -> str:
    """First 16 hex digits of a SHA-1 over the format version, the geometry (float64
    vertices, int64 faces), the grid parameters, the builder and, for SDFs, the sign
    probe (the keying of mesh_sdf.py:496-530, plus the builder, since the host and HIP
    voxelisers produce different fields)."""
    h = hashlib.sha1()
    for arr in (np.array([version], dtype=np.int64),
                np.ascontiguousarray(verts, dtype=np.float64),
                np.ascontiguousarray(faces, dtype=np.int64),
                np.asarray(params, dtype=np.float64)):
        h.update(arr.tobytes())
    h.update(device.encode())
    if probe is not None:
        h.update(np.ascontiguousarray(probe, dtype=np.float64).tobytes())
    return h.hexdigest()[:16]


def build_sdf_cached(verts, faces, res: int = 64, margin_cells: float = 4.0, cache_dir=None,
                     interior_probe=None, device: str = "cpu") -> SDFData:
    """build_sdf with a versioned on-disk cache keyed by geometry, grid, builder and sign
    probe (build_sdf_cached, mesh_sdf.py:553-563); cache_dir=None builds without caching."""
    if cache_dir is None:
        return build_sdf(verts, faces, res, margin_cells, interior_probe, device=device)
    v64 = np.asarray(verts, dtype=np.float64)
    probe = 0.5 * (v64.min(axis=0) + v64.max(axis=0)) if interior_probe is None else \
        np.asarray(interior_probe, dtype=np.float64)
    if probe.shape != (3,) or not np.isfinite(probe).all():
        raise ValueError(f"interior_probe must be a finite 3-vector, got {probe}")
    key = _field_key(_SDF_CACHE_VERSION, v64, faces, [res, margin_cells], device, probe)
    path = Path(cache_dir) / f"sdf_{key}.npz"
    if path.exists():
        return load_sdf(path)
    sdf = build_sdf(verts, faces, res, margin_cells, interior_probe, device=device)
    save_sdf(sdf, path)
    return sdf


def build_surface_cdf_cached(verts, faces, res: int = 64, band_cells: float = 3.0,
                             margin_cells: float = 4.0, cache_dir=None, device: str = "cpu") -> CDFData:
    """build_surface_cdf with a versioned on-disk cache keyed by geometry, grid and builder
    (build_surface_cdf_cached, mesh_sdf.py:496-514); cache_dir=None builds without
    caching."""
    if cache_dir is None:
        return build_surface_cdf(verts, faces, res, band_cells, margin_cells, device=device)
    key = _field_key(_CDF_CACHE_VERSION, verts, faces, [res, band_cells, margin_cells], device)
    path = Path(cache_dir) / f"cdf_{key}.npz"
    if path.exists():
        return load_cdf(path)
    cdf = build_surface_cdf(verts, faces, res, band_cells, margin_cells, device=device)
    save_cdf(cdf, path)
    return cdf


__all__ = ["SDFData", "sdf_from_function", "sphere_sdf", "CDFData", "build_surface_cdf",
           "square_sheet", "revolve_profile_open", "open_edges", "save_cdf", "load_cdf",
           "build_sdf", "winding_number", "icosphere", "save_sdf", "load_sdf", "build_sdf_cached",
           "build_surface_cdf_cached"]
