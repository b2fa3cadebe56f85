"""CPIC thin-boundary (CDF) colliders of one HipSimulator: device tag grids, the lane
table the particle pass reads, the reaction accumulators, and the host-side restamp
logic run once per substep.

Behaviour follows the reference's collider registry for CDF surfaces
(src/warpmpm/kernels/mpm_solver_warp.py:2819-3155):

* a lane restamps only when its pose changed (or its start / end window opened or
  closed); the zero / stamp / copy / mask launches cover the union of the dirty lanes'
  current boxes and the boxes their previous stamps covered, so resting surfaces cost
  nothing per substep;
* prev / cur discipline: the gather half of a pass reads tag_prev and the lanes' *_prev
  pose, the scatter half tag and the current pose. In the fused tick (lag_prev) the
  stale part of prev is brought up to cur before the next stamp overwrites cur and the
  lane pose is copied to prev before it advances; in the split substep prev is made
  equal to cur right after stamping and after the pose advances;
* poses advance once per substep like every other moving collider (center += dt v,
  quaternion exp-map step), in float64 and stored as float32.

Every launch goes to the simulator's stream; lane values reach the device through
mpm_cdf_set_lanes (the struct is passed by value into a one-workgroup copy kernel), so
nothing here synchronises with the host.
"""
from __future__ import annotations

import ctypes as C
import warnings

import numpy as np

from . import native
from .colliders import _omega_step_quat as omega_step_quat

f32 = np.float32


def _box_union(boxes):
    lo = tuple(min(b[0][i] for b in boxes) for i in range(3))
    hi = tuple(max(b[0][i] + b[1][i] for b in boxes) for i in range(3))
    return (lo, tuple(hi[i] - lo[i] for i in range(3)))


def _box_intersect(a, b):
    if a is None or b is None:
        return None
    lo = tuple(max(a[0][i], b[0][i]) for i in range(3))
    hi = tuple(min(a[0][i] + a[1][i], b[0][i] + b[1][i]) for i in range(3))
    if any(hi[i] <= lo[i] for i in range(3)):
        return None
    return (lo, tuple(hi[i] - lo[i] for i in range(3)))


def _quat_matrix(q):
    qx, qy, qz, qw = (float(v) for v in q)
    return np.array([
        [1 - 2 * (qy * qy + qz * qz), 2 * (qx * qy - qz * qw), 2 * (qx * qz + qy * qw)],
        [2 * (qx * qy + qz * qw), 1 - 2 * (qx * qx + qz * qz), 2 * (qy * qz - qx * qw)],
        [2 * (qx * qz - qy * qw), 2 * (qy * qz + qx * qw), 1 - 2 * (qx * qx + qy * qy)]])


class _Lane:
    """One registered surface: its voxel field on the device and its live pose."""

    def __init__(self, index, val, valid, res, origin, cell, center, quat, velocity, omega,
                 band, surface, friction, start_time, end_time, corners, r_max):
        self.index = index
        self.val, self.valid = val, valid
        self.res, self.origin, self.cell = res, origin, cell
        self.center = [float(f32(v)) for v in center]
        self.quat = [float(f32(v)) for v in quat]
        self.velocity = [float(f32(v)) for v in velocity]
        self.omega = [float(f32(v)) for v in omega]
        self.band = float(f32(band))
        self.surface, self.friction = surface, float(f32(friction))
        self.start_time, self.end_time = float(f32(start_time)), float(f32(end_time))
        self.corners, self.r_max = corners, r_max
        self.warned = False
        self.dirty = True
        self.active = False
        self.prev_box = None

    def is_active(self, t):
        return self.start_time <= t < self.end_time

    def moving(self):
        return any(v != 0.0 for v in self.velocity) or any(w != 0.0 for w in self.omega)


class CdfColliders:
    def __init__(self, sim):
        torch = sim.torch
        dev = sim.device
        if sim.n_scenes != 1:
            raise NotImplementedError("CDF colliders run on single-scene simulators only")
        nodes = sim.n_blocks * 64
        self.sim = sim
        self.tag = torch.zeros(nodes, dtype=torch.int32, device=dev)
        self.tag_prev = torch.zeros(nodes, dtype=torch.int32, device=dev)
        self.dist = torch.full((nodes,), 1.0e9, dtype=torch.float32, device=dev)
        self.dist_prev = torch.full((nodes,), 1.0e9, dtype=torch.float32, device=dev)
        self.block = torch.zeros(sim.n_blocks, dtype=torch.int32, device=dev)
        self.lanes_dev = torch.zeros(native.CDF_LANES_DTYPE.itemsize, dtype=torch.uint8, device=dev)
        self.reaction = torch.zeros(6 * native.MAX_CDF, dtype=torch.float32, device=dev)
        self.host = np.zeros(1, dtype=native.CDF_LANES_DTYPE)
        self.lanes: list[_Lane] = []
        self.mask_off = 0
        self.lane_prev_stale = False
        self.tag_prev_stale_box = None
        self.stamp_count = 0
        self.force_stamp = False

    @property
    def n(self) -> int:
        return len(self.lanes)

    def fill(self, G):
        c = G.cdf
        c.n_cdf = self.n
        c.mask_off = int(self.mask_off)
        c.tag, c.tag_prev = native.ptr(self.tag), native.ptr(self.tag_prev)
        c.dist, c.dist_prev = native.ptr(self.dist), native.ptr(self.dist_prev)
        c.block, c.lanes = native.ptr(self.block), native.ptr(self.lanes_dev)
        c.reaction = native.ptr(self.reaction)

    # ------------------------------------------------------------------ registry
    def add(self, values, valid, origin, cell, built_band, center, quat, velocity, omega, band,
            surface, friction, start_time, end_time, time) -> int:
        """mpm_solver_warp.py:2819-2919 (guards, AABB with a 2-node halo)."""
        torch = self.sim.torch
        lane = self.n
        if lane >= native.MAX_CDF:
            raise ValueError(f"at most {native.MAX_CDF} CDF colliders per scene")
        dx = self.sim.dx
        res = int(values.shape[0])
        if band is None:
            band = min(float(built_band), 2.0 * float(dx))
        if band > float(built_band) + 1e-12:
            raise ValueError(
                f"runtime band ({band * 1e3:.1f} mm) exceeds the band the field was built with "
                f"({float(built_band) * 1e3:.1f} mm); rebuild the CDF with a larger band_cells.")
        if band < 1.5 * float(dx):
            warnings.warn(
                f"CDF band ({band * 1e3:.1f} mm) is below 1.5 dx ({1.5 * dx * 1e3:.1f} mm): "
                f"straddling particle-node pairs can go untagged and the wall can leak.",
                RuntimeWarning, stacklevel=3)
        surface_id = {"sticky": 0, "slip": 1, "separable": 2}[surface]
        ext = (res - 1) * float(cell)
        corners = np.array([[float(origin[0]) + a * ext, float(origin[1]) + b * ext,
                             float(origin[2]) + c * ext]
                            for a in (0, 1) for b in (0, 1) for c in (0, 1)])
        r_max = float(np.linalg.norm(corners, axis=1).max())
        dev = self.sim.device
        val = torch.from_numpy(np.ascontiguousarray(values, dtype=f32).reshape(-1).copy()).to(dev)
        ok = torch.from_numpy(np.ascontiguousarray(valid, dtype=f32).reshape(-1).copy()).to(dev)
        L = _Lane(lane, val, ok, res, [float(f32(v)) for v in origin], float(f32(cell)), center,
                  quat, velocity, omega, band, surface_id, friction, start_time, end_time,
                  corners, r_max)
        L.active = L.is_active(time)
        self.lanes.append(L)
        self._sync(lane)
        return lane

    def aabb(self, L: _Lane):
        R = _quat_matrix(L.quat)
        pts = np.array(L.center) + L.corners @ R.T
        return self.sim._grid_box(pts.min(axis=0), pts.max(axis=0), halo=2)

    def _sync(self, lane):
        """Write one lane's current pose / contact state into the host mirror and upload
        the table (mpm_solver_warp.py:2921-2943)."""
        L, h = self.lanes[lane], self.host[0]
        h["center"][lane] = L.center
        h["velocity"][lane] = L.velocity
        h["omega"][lane] = L.omega
        h["friction"][lane] = L.friction
        h["type"][lane] = L.surface
        self._upload()

    def _copy_lane_prev(self):
        """mpm_solver_warp.py:3100-3105."""
        h = self.host[0]
        h["center_prev"][:] = h["center"]
        h["velocity_prev"][:] = h["velocity"]
        h["omega_prev"][:] = h["omega"]
        self.lane_prev_stale = False
        self._upload()

    def _upload(self):
        sim = self.sim
        buf = (C.c_uint8 * native.CDF_LANES_DTYPE.itemsize).from_buffer(self.host)
        native.check(sim.lib.mpm_cdf_set_lanes(sim.pG, C.addressof(buf), sim.stream()), "cdf_set_lanes")

    def set_pose(self, lane, center=None, quat=None, velocity=None, omega=None):
        """mpm_solver_warp.py:2945-2996 (teleport warning; only a changed center or quat
        restamps, velocity / omega reach the ghost through the lane table)."""
        L = self.lanes[lane]
        if not L.warned:
            jump = 0.0
            if center is not None:
                jump += float(np.linalg.norm(np.array([float(center[i]) - L.center[i] for i in range(3)])))
            if quat is not None:
                dot = abs(sum(float(quat[i]) * L.quat[i] for i in range(4)))
                jump += 2.0 * float(np.arccos(min(1.0, dot))) * L.r_max
            if jump > L.band:
                L.warned = True
                warnings.warn(
                    f"set_cdf_pose jumped the surface by up to {jump * 1e3:.2f} mm in one tick, more "
                    f"than the band ({L.band * 1e3:.2f} mm): material in the swept region can end "
                    f"up on the wrong side. Drive the pose in smaller per-tick steps.",
                    RuntimeWarning, stacklevel=3)
        moved = synced = False
        if center is not None:
            new = [float(f32(v)) for v in center]
            moved |= new != L.center
            L.center = new
        if quat is not None:
            new = [float(f32(v)) for v in quat]
            moved |= new != L.quat
            L.quat = new
        if velocity is not None:
            new = [float(f32(v)) for v in velocity]
            synced |= new != L.velocity
            L.velocity = new
        if omega is not None:
            new = [float(f32(v)) for v in omega]
            synced |= new != L.omega
            L.omega = new
        if moved:
            L.dirty = True
        if moved or synced:
            self._sync(lane)
            self.lane_prev_stale = True

    # ------------------------------------------------------------------ per substep
    def quiet(self, times) -> bool:
        """True when the per-substep logic would launch nothing during a tick whose
        substeps start at `times`: nothing stale or dirty, no lane moves, no window
        opens or closes. Such a tick may replay a captured graph."""
        if self.tag_prev_stale_box is not None or self.lane_prev_stale or self.force_stamp:
            return False
        for L in self.lanes:
            if L.dirty:
                return False
            for t in times:
                a = L.is_active(float(t))
                if a != L.active or (a and L.moving()):
                    return False
        return True

    def _dirty_box(self):
        """mpm_solver_warp.py:2998-3018: dirty lanes' boxes now and at their last stamp."""
        boxes = []
        for L in self.lanes:
            if not L.dirty:
                continue
            b = self.aabb(L)
            if b is not None:
                boxes.append(b)
            if L.prev_box is not None:
                boxes.append(L.prev_box)
            L.prev_box = b
        return _box_union(boxes) if boxes else None

    def _launch_box(self, fn, box, what):
        lo, dims = box
        lo_c = (C.c_int32 * 3)(*lo)
        hi_c = (C.c_int32 * 3)(*[lo[i] + dims[i] for i in range(3)])
        native.check(fn(self.sim.pG, lo_c, hi_c, self.sim.stream()), what)

    def reconcile_prev(self):
        """mpm_solver_warp.py:3030-3038."""
        if self.tag_prev_stale_box is not None:
            self._launch_box(self.sim.lib.mpm_cdf_copy_prev, self.tag_prev_stale_box, "cdf_copy_prev")
            self.tag_prev_stale_box = None

    def stamp(self, time: float, lag_prev: bool):
        """mpm_solver_warp.py:3040-3098: restamp the dirty lanes at `time`."""
        if self.n == 0:
            return
        self.reconcile_prev()
        for L in self.lanes:
            a = L.is_active(time)
            if a != L.active:
                L.active = a
                L.dirty = True
        if self.force_stamp:
            for L in self.lanes:
                L.dirty = True
        if not any(L.dirty for L in self.lanes):
            return
        box = self._dirty_box()
        for L in self.lanes:
            L.dirty = False
        if box is None:
            return
        self.stamp_count += 1
        sim, lib = self.sim, self.sim.lib
        self._launch_box(lib.mpm_cdf_zero_tags, box, "cdf_zero_tags")
        for L in self.lanes:
            clip = _box_intersect(box, self.aabb(L))
            if clip is None:
                continue
            p = native.CdfCollider()
            p.val, p.valid = native.ptr(L.val), native.ptr(L.valid)
            p.res, p.lane = L.res, L.index
            p.origin[:] = L.origin
            p.cell = L.cell
            p.center[:] = L.center
            p.quat[:] = L.quat
            p.band = L.band
            p.time = float(f32(time))
            p.start_time, p.end_time = L.start_time, L.end_time
            lo, dims = clip
            lo_c = (C.c_int32 * 3)(*lo)
            hi_c = (C.c_int32 * 3)(*[lo[i] + dims[i] for i in range(3)])
            native.check(lib.mpm_cdf_stamp(sim.pG, C.byref(p), lo_c, hi_c, sim.stream()), "cdf_stamp")
        if lag_prev:
            self.tag_prev_stale_box = box
        else:
            self._launch_box(lib.mpm_cdf_copy_prev, box, "cdf_copy_prev")
        lo, dims = box
        blo = tuple(lo[i] >> 2 for i in range(3))
        bhi = tuple(((lo[i] + dims[i] - 1) >> 2) + 1 for i in range(3))
        native.check(lib.mpm_cdf_block_mask(sim.pG, (C.c_int32 * 3)(*blo), (C.c_int32 * 3)(*bhi),
                                            sim.stream()), "cdf_block_mask")

    def integrate(self, time: float, dt: float, lag_prev: bool):
        """mpm_solver_warp.py:3107-3148: advance the moving active lanes one substep."""
        if self.n == 0:
            return
        if lag_prev and self.lane_prev_stale:
            self._copy_lane_prev()
        for L in self.lanes:
            if not L.is_active(time) or not L.moving():
                continue
            v, w = L.velocity, np.array(L.omega, dtype=float)
            speed = float(np.hypot(np.hypot(v[0], v[1]), v[2])) + float(np.linalg.norm(w)) * L.r_max
            if speed * dt > L.band and not L.warned:
                L.warned = True
                warnings.warn(
                    f"CDF collider surface sweeps {speed * dt * 1e3:.2f} mm per substep, more than "
                    f"the band ({L.band * 1e3:.2f} mm): material can end up on the wrong side. "
                    f"Reduce dt or the pose rate, or rebuild with a larger band.",
                    RuntimeWarning, stacklevel=4)
            L.center = [float(f32(L.center[i] + dt * v[i])) for i in range(3)]
            q = omega_step_quat(np.array(L.quat, dtype=float), w, dt)
            L.quat = [float(f32(t)) for t in q]
            self._sync(L.index)
            L.dirty = True
            self.lane_prev_stale = True
        if not lag_prev and self.lane_prev_stale:
            self._copy_lane_prev()

    def wrench(self, lane: int):
        r = self.reaction.view(native.MAX_CDF, 6)[lane].cpu().numpy().astype(np.float64)
        return r[:3], r[3:]

    def reset_wrench(self):
        self.reaction.zero_()
