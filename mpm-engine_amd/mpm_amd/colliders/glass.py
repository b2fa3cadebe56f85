"""Analytic open-top glass (solid of revolution) for the revolved-SDF cup collider.

Host-side geometry with the reference's interface (src/warpmpm/colliders/glass.py):
``GlassProfile`` (outer capped cylinder minus a fillet-dilated cavity), the closed-form
SDF ``glass_sdf_local`` whose device twin is the CUP collider kernel, cavity / solid
masks for audits, wxyz quaternion helpers, and ``cup_fill``, the jittered lattice that
fills the cavity (used to build the pour workload).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


def quat_to_mat(q) -> np.ndarray:
    """Rotation matrix (float64) of a wxyz quaternion."""
    w, x, y, z = (float(c) for c in q)
    xx, yy, zz = x * x, y * y, z * z
    return np.array([[1.0 - 2.0 * (yy + zz), 2.0 * (x * y - z * w), 2.0 * (x * z + y * w)],
                     [2.0 * (x * y + z * w), 1.0 - 2.0 * (xx + zz), 2.0 * (y * z - x * w)],
                     [2.0 * (x * z - y * w), 2.0 * (y * z + x * w), 1.0 - 2.0 * (xx + yy)]])


def quat_mul(a, b) -> np.ndarray:
    """Hamilton product of wxyz quaternions."""
    a0, a1, a2, a3 = (float(c) for c in a)
    b0, b1, b2, b3 = (float(c) for c in b)
    return np.array([a0 * b0 - a1 * b1 - a2 * b2 - a3 * b3,
                     a0 * b1 + a1 * b0 + a2 * b3 - a3 * b2,
                     a0 * b2 - a1 * b3 + a2 * b0 + a3 * b1,
                     a0 * b3 + a1 * b2 - a2 * b1 + a3 * b0])


def quat_conj(q) -> np.ndarray:
    q = np.asarray(q, dtype=np.float64)
    return np.array([q[0], -q[1], -q[2], -q[3]])


def quat_from_axis_angle(axis, angle: float) -> np.ndarray:
    axis = np.asarray(axis, dtype=np.float64)
    axis = axis / max(float(np.linalg.norm(axis)), 1e-12)
    h = 0.5 * float(angle)
    return np.concatenate([[np.cos(h)], np.sin(h) * axis])


def angular_velocity_between(q0, q1, dt: float) -> np.ndarray:
    """World angular velocity carrying wxyz q0 to q1 in time dt."""
    d = quat_mul(np.asarray(q1, dtype=np.float64), quat_conj(q0))
    d = d / np.linalg.norm(d)
    if d[0] < 0.0:
        d = -d
    s = float(np.linalg.norm(d[1:]))
    if s < 1e-9:
        return np.zeros(3)
    return d[1:] / s * (2.0 * np.arctan2(s, float(d[0])) / dt)


def world_to_local(points, pos, quat) -> np.ndarray:
    return (np.asarray(points, dtype=np.float64) - np.asarray(pos, dtype=np.float64)) @ quat_to_mat(quat)


@dataclass(frozen=True)
class GlassProfile:
    """Open-top glass, metres. Local frame: z up, origin at mid-height."""

    outer_radius: float = 0.089
    inner_radius: float = 0.069
    height: float = 0.24
    base_thickness: float = 0.050
    fillet_radius: float = 0.012

    def __post_init__(self):
        if not 0.0 < self.inner_radius < self.outer_radius:
            raise ValueError("need 0 < inner_radius < outer_radius")
        if not 0.0 < self.base_thickness < self.height:
            raise ValueError("need 0 < base_thickness < height")
        if not 0.0 <= self.fillet_radius < self.inner_radius:
            raise ValueError("need 0 <= fillet_radius < inner_radius")

    @property
    def half_height(self) -> float:
        return 0.5 * self.height

    @property
    def inner_floor_z(self) -> float:
        return self.base_thickness - 0.5 * self.height

    @property
    def rim_z(self) -> float:
        return 0.5 * self.height

    @property
    def wall_thickness(self) -> float:
        return self.outer_radius - self.inner_radius

    def inner_radius_at_z(self, z) -> np.ndarray:
        """Cavity radius at local height z (quarter-circle fillet at the floor edge)."""
        z = np.asarray(z, dtype=np.float64)
        r = np.full(z.shape, self.inner_radius)
        fr = self.fillet_radius
        if fr > 0.0:
            zf = self.inner_floor_z
            inside = (z >= zf) & (z < zf + fr)
            dz = z - (zf + fr)
            r = np.where(inside, self.inner_radius - fr + np.sqrt(np.maximum(fr * fr - dz * dz, 0.0)), r)
            r = np.where(z < zf, self.inner_radius - fr, r)
        return r


def glass_cavity_profile(profile: GlassProfile | None = None) -> np.ndarray:
    """Open (r, z) polyline of the glass cavity (floor disk, then the inner wall up to the
    rim), the CPIC contact sheet of the glass once revolved (colliders/glass.py:163-180 of
    the reference): tracing the cavity rather than the wall's mid-surface keeps CDF and
    SDF contact at the same place. The floor fillet is not represented."""
    p = GlassProfile() if profile is None else profile
    return np.array([(0.0, p.inner_floor_z), (p.inner_radius, p.inner_floor_z), (p.inner_radius, p.rim_z)])


def glass_cdf(profile: GlassProfile, dx: float, res: int = 96, n_theta: int = 96):
    """CDFData of the glass cavity sheet, built band wide enough for a 2 dx runtime band
    (examples/pour_franka.py:249-263 of the reference)."""
    from ..geometry import build_surface_cdf, revolve_profile_open

    verts, faces = revolve_profile_open(glass_cavity_profile(profile), n_theta=n_theta)
    span = float((verts.max(0) - verts.min(0)).max())
    cell = span / (res - 1 - 8)
    band_cells = max(3.0, float(np.ceil(2.2 * dx / cell)))
    return build_surface_cdf(verts, faces, res=res, band_cells=band_cells)


def _capped_cylinder(r, z, radius, z0, z1):
    zc, hh = 0.5 * (z0 + z1), 0.5 * (z1 - z0)
    dr = r - radius
    dz = np.abs(z - zc) - hh
    return np.hypot(np.maximum(dr, 0.0), np.maximum(dz, 0.0)) + np.minimum(np.maximum(dr, dz), 0.0)


def glass_sdf_local(points_local, profile: GlassProfile):
    """Signed distance (negative inside the glass material) in the local frame: outer
    capped cylinder minus the cavity, the cavity being the fillet-radius dilation of a
    capped cylinder extended above the rim (open top)."""
    p = np.asarray(points_local, dtype=np.float64)
    r = np.hypot(p[..., 0], p[..., 1])
    z = p[..., 2]
    fr = profile.fillet_radius
    outer = _capped_cylinder(r, z, profile.outer_radius, -profile.half_height, profile.half_height)
    cavity = _capped_cylinder(r, z, profile.inner_radius - fr, profile.inner_floor_z + fr,
                              profile.half_height + profile.outer_radius) - fr
    return np.maximum(outer, -cavity)


def glass_sdf(points_world, pos, quat, profile: GlassProfile):
    return glass_sdf_local(world_to_local(points_world, pos, quat), profile)


def cavity_mask(points_world, pos, quat, profile: GlassProfile, pad: float = 0.0,
                brim_clearance: float = 0.0) -> np.ndarray:
    """Points inside the open cavity below the rim."""
    loc = world_to_local(points_world, pos, quat)
    r = np.hypot(loc[:, 0], loc[:, 1])
    z = loc[:, 2]
    return ((r < profile.inner_radius_at_z(z) + pad) & (z >= profile.inner_floor_z - 1e-3)
            & (z <= profile.rim_z - brim_clearance))


def solid_mask(points_world, pos, quat, profile: GlassProfile, tol: float = 0.0) -> np.ndarray:
    """Points embedded in the glass material deeper than tol."""
    return glass_sdf(points_world, pos, quat, profile) < -tol


def cup_fill(profile: GlassProfile, h: float, fill_fraction: float = 0.80,
             clearance: float | None = None, floor_clearance: float | None = None,
             brim_clearance: float = 0.006, seed: int = 0):
    """Jittered lattice (spacing h, jitter +-h/4) filling the cavity to fill_fraction of its
    usable height. Returns (pos_local[N, 3], vol[N]) as float32."""
    clearance = h if clearance is None else clearance
    floor_clearance = h if floor_clearance is None else floor_clearance
    usable = profile.rim_z - profile.inner_floor_z - brim_clearance - floor_clearance
    top = fill_fraction * usable
    if top <= 0:
        raise ValueError("fill height is non-positive; check clearances")
    z0 = profile.inner_floor_z + floor_clearance
    rmax = profile.inner_radius - clearance
    xs = np.arange(-rmax + 0.5 * h, rmax, h)
    zs = np.arange(z0 + 0.5 * h, z0 + top, h)
    g = np.stack(np.meshgrid(xs, xs, zs, indexing="ij"), axis=-1).reshape(-1, 3)
    g = g + np.random.default_rng(seed).uniform(-0.25 * h, 0.25 * h, size=g.shape)
    keep = ((np.hypot(g[:, 0], g[:, 1]) < profile.inner_radius_at_z(g[:, 2]) - clearance)
            & (g[:, 2] >= z0) & (g[:, 2] <= z0 + top))
    pos = g[keep].astype(np.float32)
    return pos, np.full(len(pos), h ** 3, dtype=np.float32)


__all__ = ["GlassProfile", "angular_velocity_between", "cavity_mask", "cup_fill", "glass_sdf",
           "glass_sdf_local", "quat_conj", "quat_from_axis_angle", "quat_mul", "quat_to_mat",
           "solid_mask", "world_to_local"]
