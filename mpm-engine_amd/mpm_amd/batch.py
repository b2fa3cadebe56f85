"""Batched scenes: many independent MPM scenes of one grid shape advanced together.

The reference runs one scene per ``Solver`` (src/warpmpm/core/solver.py:47-642); a sweep
over hundreds of small scenes is a loop of solves whose launches each fill a small part
of the GPU. ``BatchSolver`` keeps every scene's particles in one particle array (scene
id per particle, scenes contiguous) and every scene's grid in one block-code space
(scene s owns block codes [s * B, (s + 1) * B)), so each phase of a substep is one
launch over all scenes: the fused particle pass (block-sorted chunks never mix scenes
inside an LDS tile), the grid update over one shared list of active blocks, and the
per-scene collider rows. With sparse stepping the node storage is a shared block pool:
at every tick's rebuild the active blocks of all scenes get consecutive slots
(mpm_pool_assign), so the grid takes memory in proportion to the active blocks, not to
n_scenes * n^3. A tick of `substeps` substeps is captured once as a hipGraph and
replayed.

Per-scene parameters: material type parameters are shared across scenes (per-type
arrays, as in the reference's model); yield stress, mass and the Lame parameters are
per particle, and each scene has its own colliders (``add_box(scene, ...)``).

``autotune`` picks the particle-pass launch shape (threads per workgroup, particles per
thread, or the paired pass: two particles per lane over chunks of up to 512 with one
tile and one flush for both, deterministic scatter only) and the sort interval by timing
candidates on the loaded batch from a snapshot of the state, and restores the snapshot
afterwards.
"""
from __future__ import annotations

import math
from typing import Sequence

import numpy as np

from .core import native
from .core.colliders import Collider, SceneColliders
from .core.solver import GridConfig, HipSimulator, _resolve_device

f32 = np.float32

# candidate particle-pass launch shapes: (threads per workgroup, particles per thread,
# paired pass)
LAUNCH_SHAPES = ((256, 1, False), (256, 2, False), (256, 4, False), (128, 1, False), (128, 2, False),
                 (256, 1, True))


def launch_bits(wg: int, ppt: int) -> int:
    bits = native.PASS_WG128 if wg == 128 else 0
    if ppt == 2:
        bits |= native.PASS_PPT2
    elif ppt == 4:
        bits |= native.PASS_PPT4
    elif ppt != 1:
        raise ValueError("particles per thread must be 1, 2 or 4")
    if wg not in (128, 256):
        raise ValueError("threads per workgroup must be 128 or 256")
    return bits


class BatchSolver:
    """Independent scenes sharing one grid shape, stepped together on one GPU."""

    _STATE = ("x", "v", "F", "F_trial", "C", "stress", "L", "vol", "mass", "density", "mu",
              "lam", "yield_stress", "Jp", "material", "selection", "scene")

    def __init__(self, grid: GridConfig | None = None, device: str = "auto", sparse: bool = True,
                 fused: bool = True, sort_interval: int = 1, guard_interval: int = 1,
                 wrench_slots_per_scene: int = 2, deterministic: bool = True, pool: bool = True):
        self.grid = grid or GridConfig()
        # sparse storage: node arrays for the active blocks of all scenes only (with
        # sparse stepping), instead of a dense n^3 grid per scene
        self.pool = bool(pool) and bool(sparse)
        # 64-bit fixed-point scatter and wrench sums (bit-reproducible, and faster than
        # float atomics here); False = float atomics
        self.deterministic = bool(deterministic)
        self.device = device
        self.sparse = sparse
        self.fused = fused
        self.sort_interval = int(sort_interval)
        self.guard_interval = int(guard_interval)
        self.slots_per_scene = int(wrench_slots_per_scene)
        self.launch = (256, 1, False)
        self._sim: HipSimulator | None = None
        self._tick = 0
        self._step = 0

    # ------------------------------------------------------------------ setup
    def load_scenes(self, positions: Sequence[np.ndarray], volumes: Sequence[np.ndarray]) -> "BatchSolver":
        """One (N_s, 3) position array and (N_s,) volume array per scene, in world
        coordinates of the common grid."""
        if len(positions) != len(volumes) or not positions:
            raise ValueError("need one volume array per position array, and at least one scene")
        self.device = _resolve_device(self.device)
        counts = [len(p) for p in positions]
        if min(counts) < 1:
            raise ValueError("every scene needs at least one particle")
        pos = np.concatenate([np.asarray(p, np.float32).reshape(-1, 3) for p in positions])
        vol = np.concatenate([np.asarray(v, np.float32).reshape(-1) for v in volumes])
        scene = np.repeat(np.arange(len(counts), dtype=np.int32), counts)
        self.counts = np.asarray(counts, dtype=np.int64)
        self.offsets = np.concatenate([[0], np.cumsum(self.counts)])
        self._sim = HipSimulator(pos, vol, self.grid.n_grid, self.grid.grid_lim, self.device,
                                 scene=scene, n_scenes=len(counts))
        sim = self._sim
        sim.scene_colliders = SceneColliders(len(counts), self.grid.n_grid, sim.dx,
                                             self.slots_per_scene)
        if self.pool:
            sim.enable_pool()
        if self.deterministic:
            sim.enable_deterministic()
        torch = sim.torch
        # load index of the particle in each slot (the block sort permutes slots)
        self._orig = torch.arange(len(pos), dtype=torch.int64, device=sim.device)
        return self

    @property
    def n_scenes(self) -> int:
        return len(self.counts)

    @property
    def n_particles(self) -> int:
        return int(self.offsets[-1])

    def set_material(self, material, **overrides) -> "BatchSolver":
        """Material of every particle of every scene (type parameters are shared)."""
        if hasattr(material, "resolve"):
            name, params = material.resolve()
        else:
            name, params = str(material), {}
        self._sim.set_parameters_dict({"material": name, "g": [0.0, 0.0, -9.81], **params, **overrides})
        self._sim.finalize_mu_lam()
        return self

    def _per_particle(self, values) -> np.ndarray:
        """Scalars per scene or arrays per scene (load order) -> one array in load order."""
        if np.isscalar(values):
            return np.full(self.n_particles, float(values), np.float32)
        vals = list(values)
        if len(vals) != self.n_scenes:
            raise ValueError(f"expected {self.n_scenes} per-scene entries, got {len(vals)}")
        parts = []
        for s, v in enumerate(vals):
            v = np.asarray(v, np.float32).reshape(-1)
            if v.size == 1:
                v = np.full(int(self.counts[s]), float(v[0]), np.float32)
            if v.size != self.counts[s]:
                raise ValueError(f"scene {s}: {v.size} values for {self.counts[s]} particles")
            parts.append(v)
        return np.concatenate(parts)

    def _load_order_to_slots(self, a: np.ndarray):
        sim = self._sim
        t = sim.torch.from_numpy(a).to(sim.device)
        return t[self._orig]

    def set_yield_stress(self, values) -> "BatchSolver":
        """Per-particle yield stress (a scalar per scene or an array per scene)."""
        self._sim.yield_stress.copy_(self._load_order_to_slots(self._per_particle(values)))
        return self

    def set_velocity(self, values) -> "BatchSolver":
        """Per-particle initial velocity: an (N_s, 3) array (or a 3-vector) per scene."""
        parts = []
        for s, v in enumerate(values):
            v = np.asarray(v, np.float32)
            parts.append(np.broadcast_to(v, (int(self.counts[s]), 3)))
        a = np.ascontiguousarray(np.concatenate(parts))
        self._sim.v.copy_(self._load_order_to_slots(a).t())
        return self

    # ------------------------------------------------------------------ colliders
    def _set(self, scene: int):
        return self._sim.scene_colliders.sets[scene]

    def add_plane(self, point, normal, surface: str = "sticky", friction: float = 0.0) -> "BatchSolver":
        """The same plane collider in every scene (mpm_solver_warp.py:1880-2018)."""
        if surface == "sticky" and friction != 0:
            raise ValueError("friction must be 0 on sticky surfaces.")
        scale = 1.0 / float(f32(math.sqrt(float(sum(float(x) ** 2 for x in normal)))))
        nrm = [scale * float(x) for x in normal]
        axis = next((i for i in range(3) if abs(abs(nrm[i]) - 1.0) < 1e-6), -1)
        for s in range(self.n_scenes):
            c = Collider(native.COLLIDER_PLANE, "plane", 0.0, 999.0)
            c.point = [float(f32(v)) for v in point]
            c.normal = [float(f32(v)) for v in nrm]
            c.surface = {"sticky": 0, "slip": 1, "cut": 11}.get(surface, 2)
            c.friction = float(f32(friction))
            if axis >= 0:
                c.aabb_kind = "plane"
                c.plane_axis = axis
                c.plane_offset = float(point[axis])
                c.plane_sign = float(nrm[axis])
            self._set(s).add(c)
        return self

    def add_domain_walls(self, start_time: float = 0.0, end_time: float = 1.0e9) -> "BatchSolver":
        """Outward-velocity clamp within 3 cells of each face in every scene (add_bounding_box,
        mpm_solver_warp.py:2287-2380). One whole-grid walls collider per scene: the wall
        rule tests all six faces at every node it visits and is idempotent, so this is the
        same as the reference's six face shells."""
        n = self.grid.n_grid
        for s in range(self.n_scenes):
            c = Collider(native.COLLIDER_WALLS, "domain_walls", start_time, end_time)
            c.aabb_kind = "fixed"
            c.fixed_box = ((0, 0, 0), (n, n, n))
            self._set(s).add(c)
        return self

    def add_box(self, scene: int, center, half_size, velocity=(0.0, 0.0, 0.0),
                start_time: float = 0.0, end_time: float = 1.0e9) -> int:
        """Kinematic box in one scene (set_velocity_on_cuboid, mpm_solver_warp.py:
        2028-2106); returns its handle within that scene."""
        c = Collider(native.COLLIDER_BOX, "cuboid_velocity", start_time, end_time)
        c.point = [float(f32(v)) for v in center]
        c.size = [float(f32(v)) for v in half_size]
        c.velocity = [float(f32(v)) for v in velocity]
        c.aabb_kind = "box"
        return self._set(scene).add(c, wrench=True)

    def set_box(self, scene: int, handle: int, center=None, velocity=None) -> "BatchSolver":
        c = self._set(scene).items[handle]
        if center is not None:
            c.point = [float(f32(v)) for v in center]
        if velocity is not None:
            c.velocity = [float(f32(v)) for v in velocity]
        return self

    def box_center(self, scene: int, handle: int) -> np.ndarray:
        return np.asarray(self._set(scene).items[handle].point, dtype=np.float64)

    def _slot(self, scene: int, handle: int) -> int:
        c = self._set(scene).items[handle]
        if c.slot < 0:
            raise ValueError(f"collider {handle} of scene {scene} has no reaction accumulator")
        return scene * self.slots_per_scene + c.slot

    def tool_force(self, scene: int, handle: int, dt: float) -> np.ndarray:
        """Accumulated reaction impulse of a box / dt (Solver.tool_force, solver.py:415-427)."""
        k = self._slot(scene, handle)
        w = self._sim.wrench
        if 6 * k + 3 > w.numel():
            return np.zeros(3)
        return w[6 * k:6 * k + 3].cpu().numpy().astype(float) / dt

    def tool_forces(self, handle: int, dt: float) -> np.ndarray:
        """(n_scenes, 3) tool_force of the collider `handle` of every scene."""
        w = self._sim.wrench
        need = 6 * self.n_scenes * self.slots_per_scene
        out = np.zeros((self.n_scenes, 3))
        if w.numel() < need:
            return out
        slots = np.array([self._slot(s, handle) for s in range(self.n_scenes)])
        ww = w.cpu().numpy().astype(float).reshape(-1, 6)
        return ww[slots, :3] / dt

    def reset_tool_force(self, scene: int | None = None, handle: int = 0) -> "BatchSolver":
        scenes = range(self.n_scenes) if scene is None else (scene,)
        w = self._sim.wrench
        for s in scenes:
            k = self._slot(s, handle)
            if 6 * k + 6 <= w.numel():
                w[6 * k:6 * k + 6].zero_()
        return self

    # ------------------------------------------------------------------ stepping
    def set_launch(self, wg: int = 256, ppt: int = 1, paired: bool = False) -> "BatchSolver":
        """Particle-pass launch shape; paired (with wg 256, ppt 1) selects the paired
        interior pass, which the deterministic scatter runs (float atomics keep the
        single-particle pass)."""
        if paired and (wg, ppt) != (256, 1):
            raise ValueError("the paired pass runs 256-lane workgroups with one chunk each")
        self._sim.launch_bits = launch_bits(wg, ppt)
        self._sim.pair_pass = bool(paired)
        self.launch = (wg, ppt, bool(paired))
        return self

    def _sort(self) -> bool:
        moved = self._sim.sort_by_block(cells=True)
        if moved:
            self._orig = self._orig[self._sim._last_perm]
        return moved

    def _guard(self, dt: float, substeps: int) -> None:
        """Error flags, edge check and (dense mode) the launch box: the union of every
        scene's particle box, padded as in Solver._update_grid_box (solver.py:485-518)."""
        lo_x, hi_x, vmax, bad, flag = self._sim.read_bounds()
        dx, lim = self.grid.dx, self.grid.grid_lim
        if flag & native.ERR_SCENE:
            raise RuntimeError("internal: a particle carries a scene id outside the batch")
        if flag & native.ERR_SOLID_IN_VISCOUS:
            raise RuntimeError("internal: a solid material ran in the viscous-only particle pass")
        if flag & native.ERR_STENCIL:
            raise RuntimeError("a particle stencil left its scene's grid during the last tick")
        if flag & native.ERR_POOL:
            raise RuntimeError("a particle scattered into a block outside its scene's active set "
                               "during the last tick (it moved more than a block within the tick)")
        if bad or lo_x.min() < 1.5 * dx or hi_x.max() > lim - 2.5 * dx:
            raise RuntimeError(
                f"particles of some scene within 2 cells of the grid edge (x in "
                f"[{lo_x.min():.4f}, {hi_x.max():.4f}] m, domain [0, {lim}] m) or non-finite")
        pad = 3.0 * dx + 1.5 * float(vmax) * dt * substeps
        self._sim.grid_launch_box = self._sim._grid_box(lo_x - pad, hi_x + pad, halo=0)

    def step(self, dt: float, substeps: int = 1) -> "BatchSolver":
        """One control tick of `substeps` substeps for every scene: guard, optional block
        sort, active-block rebuild, then the tick (a hipGraph replay after the first)."""
        sim = self._sim
        # collider rows first: the host builds them while the GPU finishes the last tick
        sim.prepare_tick(dt, substeps)
        try:
            if self._tick % max(1, self.guard_interval) == 0:
                self._guard(dt, substeps)
            if self.sort_interval and self._tick % self.sort_interval == 0:
                self._sort()
            if self.sparse:
                sim.rebuild_active_blocks()
        except BaseException:
            sim.drop_prepared()
            raise
        self._tick += 1
        sim.run_tick(dt, substeps, fused=self.fused, sparse=self.sparse)
        self._step += substeps
        return self

    # ------------------------------------------------------------------ readouts
    def _scene_rows(self, t, scene: int) -> np.ndarray:
        """Rows of slot-ordered tensor t belonging to `scene`, in load order."""
        a, b = int(self.offsets[scene]), int(self.offsets[scene + 1])
        orig = self._orig[a:b]
        out = t.new_empty((b - a,) + tuple(t.shape[1:]))
        out[orig - a] = t[a:b]
        return out.cpu().numpy()

    def x(self, scene: int) -> np.ndarray:
        return self._scene_rows(self._sim.x.t(), scene)

    def v(self, scene: int) -> np.ndarray:
        return self._scene_rows(self._sim.v.t(), scene)

    def F(self, scene: int) -> np.ndarray:
        return self._scene_rows(self._sim.F.t(), scene).reshape(-1, 3, 3)

    def stress(self, scene: int) -> np.ndarray:
        return self._scene_rows(self._sim.stress.t(), scene).reshape(-1, 3, 3)

    def grid_dense(self):
        """Dense (n_scenes, n, n, n) mass and (n_scenes, n, n, n, 3) velocities."""
        sim = self._sim
        torch = sim.torch
        n, ns = self.grid.n_grid, self.n_scenes
        m = torch.empty(ns, n, n, n, dtype=torch.float32, device=sim.device)
        vi = torch.empty(ns, n, n, n, 3, dtype=torch.float32, device=sim.device)
        vo = torch.empty(ns, n, n, n, 3, dtype=torch.float32, device=sim.device)
        native.check(sim.lib.mpm_grid_to_dense(sim.pG, native.ptr(m), native.ptr(vi), native.ptr(vo),
                                               sim.stream()), "grid_to_dense")
        return m, vi, vo

    # ------------------------------------------------------------------ autotuning
    def _snapshot(self):
        sim = self._sim
        state = {k: getattr(sim, k).clone() for k in self._STATE}
        cols = [[(list(c.point), np.array(c.rot), list(c.quat)) for c in cs.items]
                for cs in sim.scene_colliders.sets]
        return {"state": state, "orig": self._orig.clone(), "wrench": sim.wrench.clone(),
                "time": sim.time, "cols": cols, "tick": self._tick, "step": self._step}

    def _restore(self, snap):
        sim = self._sim
        for k, t in snap["state"].items():
            getattr(sim, k).copy_(t)
        self._orig = snap["orig"].clone()
        w = snap["wrench"]
        sim.wrench.zero_()
        sim.wrench[:w.numel()] = w
        sim.time = snap["time"]
        for cs, saved in zip(sim.scene_colliders.sets, snap["cols"]):
            for c, (p, r, q) in zip(cs.items, saved):
                c.point, c.rot, c.quat = list(p), np.array(r), list(q)
        self._tick, self._step = snap["tick"], snap["step"]
        sim.node_in.zero_()
        sim.node_out.zero_()
        if sim._blk is not None:
            for k in ("cur", "prev"):
                sim._blk[k].zero_()

    def _time_ticks(self, dt, substeps, ticks) -> list[float]:
        torch = self._sim.torch
        out = []
        for _ in range(ticks):
            torch.cuda.synchronize(self._sim.device)
            a = torch.cuda.Event(enable_timing=True)
            b = torch.cuda.Event(enable_timing=True)
            a.record()
            self.step(dt, substeps)
            b.record()
            torch.cuda.synchronize(self._sim.device)
            out.append(a.elapsed_time(b) / 1e3)
        return out

    def autotune(self, dt: float, substeps: int, shapes=LAUNCH_SHAPES, sort_intervals=(1, 2, 4, 8),
                 ticks: int = 2) -> dict:
        """Choose the particle-pass launch shape and the sort interval for this batch by
        timing candidates from a snapshot of the state (restored afterwards, so the
        simulation continues from where it was; only float-atomic summation order can
        differ from an untuned run).

        Launch shape: each candidate runs one warm-up tick and `ticks` timed ticks from
        the snapshot with a sort at the first tick. Sort interval I: the average tick
        cost sort_cost / I + mean(tick time j ticks after a sort, j < I) is modelled from
        one sort's cost and max(sort_intervals) consecutive unsorted ticks."""
        sim = self._sim
        torch = sim.torch
        use_graph = sim.use_graph
        sim.use_graph = False  # eager launches: no capture cost inside the timings
        saved_si, saved_tick = self.sort_interval, self._tick
        snap = self._snapshot()
        result = {"shapes": {}, "sort": {}}
        if sim.deterministic:  # the fixed-point scatter runs 256-lane workgroups only
            shapes = tuple(sh for sh in shapes if sh[0] == 256) or ((256, 1, False),)
        else:  # the paired pass is a deterministic-scatter kernel
            shapes = tuple(sh for sh in shapes if not sh[2]) or ((256, 1, False),)
        try:
            best = None
            for wg, ppt, paired in shapes:
                self._restore(snap)
                self.set_launch(wg, ppt, paired)
                self.sort_interval = 1
                self._tick = 0
                self._time_ticks(dt, substeps, 1)
                t = float(np.mean(self._time_ticks(dt, substeps, ticks)))
                result["shapes"]["paired512" if paired else f"wg{wg}_ppt{ppt}"] = t
                if best is None or t < best[0]:
                    best = (t, wg, ppt, paired)
            self.set_launch(best[1], best[2], best[3])
            # sort interval model
            self._restore(snap)
            torch.cuda.synchronize(sim.device)
            a = torch.cuda.Event(enable_timing=True)
            b = torch.cuda.Event(enable_timing=True)
            a.record()
            self._sort()
            b.record()
            torch.cuda.synchronize(sim.device)
            t_sort = a.elapsed_time(b) / 1e3
            self.sort_interval = 0
            J = max(sort_intervals)
            t_j = self._time_ticks(dt, substeps, J)
            model = {I: t_sort / I + float(np.mean(t_j[:I])) for I in sort_intervals}
            best_I = min(model, key=model.get)
            result["sort"] = {"sort_s": t_sort, "tick_s_after_sort": t_j,
                              "modelled_tick_s": {str(k): v for k, v in model.items()}}
            self.sort_interval = int(best_I)
        except Exception:
            self.sort_interval = saved_si
            raise
        finally:
            self._restore(snap)
            self._tick = saved_tick
            sim.use_graph = use_graph
        result["chosen"] = {"threads_per_workgroup": self.launch[0],
                            "particles_per_thread": self.launch[1],
                            "paired": self.launch[2],
                            "sort_interval": self.sort_interval}
        return result
