"""Collider registry and per-tick pose schedule (host side).

The reference registers each grid-BC collider as a Warp kernel plus a host ``modify_bc``
callable that integrates its pose after every substep launch, and an AABB callable that
restricts the launch (mpm_solver_warp.py:440-449, 1283-1336, 1880-2814). Here a collider
is a typed record; before each control tick ``ColliderSet.schedule`` runs the same host
pose integration for all substeps of the tick at once and writes one ``mpm_collider_t``
row per (substep, collider) into a table the grid kernel reads from device memory. The
float64 arithmetic of the reference's integrators is kept (points advance as
``fp32(p + dt * v)``, cup rotations on a float64 shadow with an exact Rodrigues
increment, SDF quaternions by the exact exponential map and re-stored as fp32), so the
schedule reproduces the reference's per-substep poses exactly.
"""
from __future__ import annotations

import warnings

import numpy as np

from . import native

f32 = np.float32


def _f32(v) -> float:
    return float(f32(v))


def grid_box(lo_w, hi_w, dx: float, n, halo: int = 1):
    """World box -> clamped node box (lo, dims), or None if it misses the grid
    (mpm_solver_warp.py:1868-1878); dx is the fp32 model dx."""
    lo = [max(0, int(np.floor(float(lo_w[i]) / dx)) - halo) for i in range(3)]
    hi = [min(n[i], int(np.ceil(float(hi_w[i]) / dx)) + halo + 1) for i in range(3)]
    dims = tuple(hi[i] - lo[i] for i in range(3))
    if min(dims) <= 0:
        return None
    return tuple(lo), dims


def _grid_box_vec(lo_w, hi_w, dx, n, halo=1):
    """Vectorised grid_box over rows of (S, 3) float64 arrays -> (lo, hi, active)."""
    lo = np.maximum(0, np.floor(lo_w / dx).astype(np.int64) - halo)
    hi = np.minimum(np.asarray(n), np.ceil(hi_w / dx).astype(np.int64) + halo + 1)
    active = np.all(hi - lo > 0, axis=1)
    return lo, hi, active


def _omega_step_quat(q, omega, dt):
    """Exact exponential-map step of an xyzw quaternion by world angular velocity omega
    over dt (left multiplication), renormalised; identity below 1e-12 rad."""
    w = np.asarray(omega, dtype=float)
    speed = float(np.linalg.norm(w))
    if speed * dt < 1e-12:
        return q
    half = 0.5 * speed * dt
    ax = w / speed
    sh = np.sin(half)
    bx, by, bz, bw = ax[0] * sh, ax[1] * sh, ax[2] * sh, np.cos(half)
    qx, qy, qz, qw = q
    out = np.array([bw * qx + bx * qw + by * qz - bz * qy,
                    bw * qy - bx * qz + by * qw + bz * qx,
                    bw * qz + bx * qy - by * qx + bz * qw,
                    bw * qw - bx * qx - by * qy - bz * qz])
    return out / np.linalg.norm(out)


def _quat_R(q):
    qx, qy, qz, qw = q
    return np.array([
        [1 - 2 * (qy * qy + qz * qz), 2 * (qx * qy - qz * qw), 2 * (qx * qz + qy * qw)],
        [2 * (qx * qy + qz * qw), 1 - 2 * (qx * qx + qz * qz), 2 * (qy * qz - qx * qw)],
        [2 * (qx * qz - qy * qw), 2 * (qy * qz + qx * qw), 1 - 2 * (qx * qx + qy * qy)]])


class Collider:
    """One registered grid-BC collider. Pose fields hold fp32-representable floats,
    mirroring the Warp struct fields of the reference."""

    def __init__(self, kind: int, label: str, start_time: float, end_time: float):
        self.kind = kind
        self.label = label
        self.start_time = _f32(start_time)
        self.end_time = _f32(end_time)
        self.surface = 0
        self.reset = 0
        self.slot = -1
        self.point = [0.0, 0.0, 0.0]
        self.normal = [0.0, 0.0, 0.0]
        self.size = [0.0, 0.0, 0.0]
        self.velocity = [0.0, 0.0, 0.0]
        self.omega = [0.0, 0.0, 0.0]
        self.rot = np.eye(3)            # float64 shadow (cup)
        self.quat = [0.0, 0.0, 0.0, 1.0]  # xyzw (sdf)
        self.friction = 0.0
        self.sticky_depth = 0.0
        self.contact_band = 0.0
        self.band = 0.0
        self.shape = (0.0, 0.0, 0.0, 0.0, 0.0)  # cup radii / heights
        self.sdf = None                 # (val_dev, grad_dev, res, origin, cell, r_max)
        self.occupancy = None
        self.fixed_box = None           # static launch box (walls / point cloud)
        self.aabb_kind = None           # 'plane' | 'box' | 'ball' | 'sdf' | 'fixed' | None
        self.plane_axis = None
        self.plane_offset = 0.0   # float64 plane coordinate along plane_axis
        self.plane_sign = 0.0
        self.ball_r = 0.0
        self.sdf_corners = None
        self.guard = None
        self.moving = kind in (native.COLLIDER_BOX, native.COLLIDER_CUP, native.COLLIDER_SDF)

    def live(self, t: float) -> bool:
        return self.start_time <= t < self.end_time


class ColliderSet:
    def __init__(self, n_grid: int, dx: float):
        self.n = (n_grid, n_grid, n_grid)
        self.dx = dx
        self.items: list[Collider] = []
        self.n_slots = 0
        self.restrict_bc = True

    def add(self, c: Collider, wrench: bool = False) -> int:
        if wrench:
            c.slot = self.n_slots
            self.n_slots += 1
        self.items.append(c)
        return len(self.items) - 1

    def poses(self):
        """Snapshot of the pose state a schedule advances (point, rotation, quaternion)."""
        return [(list(c.point), np.array(c.rot, copy=True), list(c.quat)) for c in self.items]

    def set_poses(self, snap):
        for c, (p, r, q) in zip(self.items, snap):
            c.point, c.rot, c.quat = p, r, q

    # ------------------------------------------------------------------ boxes
    def aabb(self, c: Collider):
        """Current restricted launch box of c, None when it misses the grid, or the
        full grid when c has no box (the reference's per-collider aabb callables)."""
        full = ((0, 0, 0), self.n)
        if not self.restrict_bc or c.aabb_kind is None:
            return full
        if c.aabb_kind == "fixed":
            return c.fixed_box
        if c.aabb_kind == "plane":
            big = 1.0e9
            lo_w, hi_w = [-big] * 3, [big] * 3
            ax = c.plane_axis
            if c.plane_sign > 0.0:
                hi_w[ax] = c.plane_offset
            else:
                lo_w[ax] = c.plane_offset
            return grid_box(lo_w, hi_w, self.dx, self.n, 1)
        if c.aabb_kind == "box":
            p, s = c.point, c.size
            return grid_box([p[i] - s[i] for i in range(3)], [p[i] + s[i] for i in range(3)],
                            self.dx, self.n, 1)
        if c.aabb_kind == "ball":
            p, r = c.point, c.ball_r
            return grid_box([p[i] - r for i in range(3)], [p[i] + r for i in range(3)],
                            self.dx, self.n, 1)
        if c.aabb_kind == "sdf":
            pts = np.asarray(c.point) + c.sdf_corners @ _quat_R(c.quat).T
            return grid_box(pts.min(0), pts.max(0), self.dx, self.n, 1)
        raise AssertionError(c.aabb_kind)

    # ------------------------------------------------------------------ schedule
    @staticmethod
    def times(time0: float, dt: float, substeps: int):
        """float64 start time of every substep (time += dt per substep, as the reference's
        host loop accumulates it) and the end time."""
        times = np.empty(substeps, dtype=np.float64)
        t = float(time0)
        for s in range(substeps):
            times[s] = t
            t = t + dt
        return times, t

    def schedule(self, time0: float, dt: float, substeps: int):
        """Rows [substeps, ncoll] of mpm_collider_t for one tick starting at time0, and
        the end time. Advances every collider's pose through the tick exactly like the
        reference's per-substep modify_bc calls."""
        S, K = substeps, len(self.items)
        times, t = self.times(time0, dt, S)
        tab = np.zeros((S, K), dtype=native.COLLIDER_DTYPE)
        if K == 0:
            return tab, t
        tab["time"] = times.astype(f32)[:, None]
        tab["dt"] = dt
        for k, c in enumerate(self.items):
            self.fill(c, tab[:, k], times, dt)
        return tab, t

    def fill(self, c: Collider, col, times, dt):
        """Write collider c's rows for the substeps at `times` into col (all but the
        time / dt fields), advancing a moving collider's pose."""
        col["kind"] = c.kind
        col["surface"] = c.surface
        col["reset"] = c.reset
        col["slot"] = c.slot
        col["start_time"] = c.start_time
        col["end_time"] = c.end_time
        col["normal"] = c.normal
        col["size"] = c.size
        col["friction"] = c.friction
        col["sticky_depth"] = c.sticky_depth
        col["contact_band"] = c.contact_band
        col["band"] = c.band
        (col["outer_radius"], col["inner_radius"], col["half_height"],
         col["inner_floor_z"], col["fillet_radius"]) = c.shape
        if c.sdf is not None:
            val, grad, res, origin, cell, _ = c.sdf
            col["sdf_val"], col["sdf_grad"], col["res"] = val.data_ptr(), grad.data_ptr(), res
            col["origin"] = origin
            col["cell"] = cell
        if c.occupancy is not None:
            col["occupancy"] = c.occupancy.data_ptr()
        if c.moving:
            self._integrate(c, col, times, dt)
        else:
            col["point"] = c.point
            col["velocity"] = c.velocity
            col["omega"] = c.omega
            col["rot"] = np.asarray(c.rot, dtype=f32).ravel()
            col["quat"] = c.quat
            box = self.aabb(c)
            if box is None:
                col["active"] = 0
            else:
                col["active"] = 1
                col["lo"] = box[0]
                col["hi"] = [box[0][i] + box[1][i] for i in range(3)]

    def static_row(self, c: Collider):
        """The (substep-independent) row of a non-moving collider, built once: plane,
        wall and point-cloud colliders have no pose setters."""
        row = getattr(c, "_static_row", None)
        if row is None or c._static_restrict != self.restrict_bc:
            row = np.zeros(1, dtype=native.COLLIDER_DTYPE)
            self.fill(c, row, np.zeros(1), 0.0)
            c._static_row, c._static_restrict = row, self.restrict_bc
        return row[0]

    def _integrate(self, c: Collider, col, times, dt):
        S = len(times)
        pts = np.empty((S, 3), dtype=np.float64)
        p = list(c.point)
        v = c.velocity
        if c.kind == native.COLLIDER_CUP:
            rots = np.empty((S, 9), dtype=f32)
            w = np.array(c.omega, dtype=np.float64)
            wn = float(np.linalg.norm(w))
            ang = wn * dt
            inc = None
            if ang > 1.0e-12:
                ax = w / wn
                kx = np.array([[0.0, -ax[2], ax[1]], [ax[2], 0.0, -ax[0]], [-ax[1], ax[0], 0.0]])
                inc = np.eye(3) + np.sin(ang) * kx + (1.0 - np.cos(ang)) * (kx @ kx)
        if c.kind == native.COLLIDER_SDF:
            quats = np.empty((S, 4), dtype=np.float64)
            wv = np.array(c.omega, dtype=float)
            speed = float(np.hypot(np.hypot(v[0], v[1]), v[2])) + float(np.linalg.norm(wv)) * c.guard["r_max"]
        rot_f32 = np.asarray(c.rot, dtype=f32).ravel()
        for s in range(S):
            t = times[s]
            pts[s] = p
            if c.kind == native.COLLIDER_CUP:
                rots[s] = rot_f32
            elif c.kind == native.COLLIDER_SDF:
                quats[s] = c.quat
            if c.start_time <= t < c.end_time:
                p = [_f32(p[i] + dt * v[i]) for i in range(3)]
                if c.kind == native.COLLIDER_CUP and inc is not None:
                    c.rot = inc @ c.rot
                    rot_f32 = np.asarray(c.rot, dtype=f32).ravel()
                elif c.kind == native.COLLIDER_SDF:
                    if speed * dt > c.guard["band"] and not c.guard["warned"]:
                        c.guard["warned"] = True
                        warnings.warn(
                            f"SDF collider surface sweeps {speed * dt * 1e3:.2f} mm per substep, "
                            f"more than the contact band ({c.guard['band'] * 1e3:.2f} mm): fast "
                            f"material can tunnel through. Reduce dt, velocity/omega, or raise "
                            f"band (and the SDF margin).", RuntimeWarning, stacklevel=4)
                    q = _omega_step_quat(np.array(c.quat, dtype=float), wv, dt)
                    c.quat = [_f32(x) for x in q]
        c.point = p
        col["point"] = pts.astype(f32)
        col["velocity"] = c.velocity
        col["omega"] = c.omega
        if c.kind == native.COLLIDER_CUP:
            col["rot"] = rots
        else:
            col["rot"] = np.asarray(c.rot, dtype=f32).ravel()
        if c.kind == native.COLLIDER_SDF:
            col["quat"] = quats.astype(f32)
        else:
            col["quat"] = c.quat
        # restricted launch boxes from the per-substep poses
        ptsf = pts.astype(f32).astype(np.float64)
        if not self.restrict_bc:
            col["active"] = 1
            col["lo"] = (0, 0, 0)
            col["hi"] = self.n
            return
        if c.aabb_kind == "box":
            s = np.asarray(c.size, dtype=np.float64)
            lo, hi, act = _grid_box_vec(ptsf - s, ptsf + s, self.dx, self.n)
        elif c.aabb_kind == "ball":
            lo, hi, act = _grid_box_vec(ptsf - c.ball_r, ptsf + c.ball_r, self.dx, self.n)
        else:  # sdf: oriented corner box per substep
            lo = np.empty((S, 3), np.int64)
            hi = np.empty((S, 3), np.int64)
            act = np.empty(S, bool)
            for s in range(S):
                q = quats[s].astype(f32).astype(np.float64)
                corners = ptsf[s] + c.sdf_corners @ _quat_R(q).T
                b = grid_box(corners.min(0), corners.max(0), self.dx, self.n, 1)
                act[s] = b is not None
                if b is not None:
                    lo[s] = b[0]
                    hi[s] = [b[0][i] + b[1][i] for i in range(3)]
                else:
                    lo[s] = hi[s] = 0
        col["lo"] = lo
        col["hi"] = hi
        col["active"] = act.astype(np.int32)


def _schedule_boxes(cols, sets, blk, times, dt):
    """Vectorised pose integration of one kinematic box per scene (the same fp32
    round trip as ColliderSet._integrate: p <- fp32(p + dt v) while live) into blk
    [substeps, n_scenes]; advances every box."""
    S, ns = len(times), len(cols)
    P = np.array([c.point for c in cols], dtype=np.float64)
    V = np.array([c.velocity for c in cols], dtype=np.float64)
    st = np.array([c.start_time for c in cols], dtype=np.float64)
    en = np.array([c.end_time for c in cols], dtype=np.float64)
    pts = np.empty((S, ns, 3), dtype=np.float64)
    for s in range(S):
        pts[s] = P
        live = (st <= times[s]) & (times[s] < en)
        if live.any():
            P = np.where(live[:, None], (P + dt * V).astype(f32).astype(np.float64), P)
    for c, p in zip(cols, P):
        c.point = [float(v) for v in p]
    for fld in ("kind", "surface", "reset", "slot", "start_time", "end_time", "friction"):
        blk[fld] = np.array([getattr(c, fld) for c in cols])[None, :]
    for fld in ("normal", "size", "velocity", "omega", "quat"):
        blk[fld] = np.array([getattr(c, fld) for c in cols], dtype=f32)[None, :, :]
    blk["rot"] = np.array([np.asarray(c.rot, dtype=f32).ravel() for c in cols])[None, :, :]
    blk["point"] = pts.astype(f32)
    cs0 = sets[0]
    if not cs0.restrict_bc:
        blk["active"] = 1
        blk["lo"] = (0, 0, 0)
        blk["hi"] = cs0.n
        return
    ptsf = pts.astype(f32).astype(np.float64).reshape(-1, 3)
    size = np.broadcast_to(np.array([c.size for c in cols], dtype=np.float64)[None], (S, ns, 3)).reshape(-1, 3)
    lo, hi, act = _grid_box_vec(ptsf - size, ptsf + size, cs0.dx, cs0.n)
    blk["lo"] = lo.reshape(S, ns, 3)
    blk["hi"] = hi.reshape(S, ns, 3)
    blk["active"] = act.reshape(S, ns).astype(np.int32)


class SceneColliders:
    """One ColliderSet per batched scene and the per-tick table builder: rows
    [substeps, n_scenes * K], scene-major within a substep as the grid kernel reads them
    (K = the most colliders of any scene; shorter scenes are padded with inactive rows).
    Wrench slots become global: scene * slots_per_scene + the scene's own slot. The rows
    of non-moving colliders are built once into a template; each tick copies it and
    integrates the moving colliders (kinematic boxes vectorised over scenes)."""

    def __init__(self, n_scenes: int, n_grid: int, dx: float, slots_per_scene: int):
        self.sets = [ColliderSet(n_grid, dx) for _ in range(n_scenes)]
        self.slots_per_scene = int(slots_per_scene)
        self._template = None
        self._sig = None

    def _signature(self):
        return tuple((len(cs.items), cs.restrict_bc) for cs in self.sets)

    def poses(self):
        return [cs.poses() for cs in self.sets]

    def set_poses(self, snap):
        for cs, sn in zip(self.sets, snap):
            cs.set_poses(sn)

    def _build_template(self, K):
        ns = len(self.sets)
        tpl = np.zeros((ns, K), dtype=native.COLLIDER_DTYPE)
        tpl["slot"] = -1
        for s, cs in enumerate(self.sets):
            if cs.n_slots > self.slots_per_scene:
                raise ValueError(f"scene {s} has {cs.n_slots} wrench colliders, more than the "
                                 f"{self.slots_per_scene} slots reserved per scene")
            for k, c in enumerate(cs.items):
                if not c.moving:
                    tpl[s, k] = cs.static_row(c)
        slot = tpl["slot"]
        scene = np.arange(ns, dtype=np.int32)[:, None]
        tpl["slot"] = np.where(slot >= 0, slot + scene * self.slots_per_scene, -1)
        return tpl

    def schedule(self, time0: float, dt: float, substeps: int):
        """(table, end time, K, total wrench slots) for one tick starting at time0."""
        ns = len(self.sets)
        K = max((len(cs.items) for cs in self.sets), default=0)
        times, t_end = ColliderSet.times(time0, dt, substeps)
        sig = self._signature()
        if self._template is None or self._sig != sig:
            self._template, self._sig = self._build_template(K), sig
        tab = np.empty((substeps, ns, K), dtype=native.COLLIDER_DTYPE)
        tab[:] = self._template[None]
        if K == 0:
            return tab.reshape(substeps, 0), t_end, 0, ns * self.slots_per_scene
        same_restrict = len({cs.restrict_bc for cs in self.sets}) <= 1
        scene_slot = np.arange(ns, dtype=np.int32) * self.slots_per_scene
        for k in range(K):
            cols = [cs.items[k] if k < len(cs.items) else None for cs in self.sets]
            if not any(c is not None and c.moving for c in cols):
                continue
            blk = tab[:, :, k]
            if same_restrict and all(c is not None and c.kind == native.COLLIDER_BOX and
                                     c.aabb_kind == "box" for c in cols):
                _schedule_boxes(cols, self.sets, blk, times, dt)
            else:
                for s, c in enumerate(cols):
                    if c is not None and c.moving:
                        self.sets[s].fill(c, blk[:, s], times, dt)
            slot = blk["slot"]
            blk["slot"] = np.where(slot >= 0, slot + scene_slot[None, :], -1)
        tab["time"] = times.astype(f32)[:, None, None]
        tab["dt"] = dt
        return tab.reshape(substeps, ns * K), t_end, K, ns * self.slots_per_scene
