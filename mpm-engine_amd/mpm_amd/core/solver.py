"""Typed solver API on MI355X: the drop-in for warpmpm's ``Solver`` (src/warpmpm/core/
solver.py:47-642) with ``device="hip"``.

Same constructor flags (grid, device, guard_interval, sparse, fused, sort_interval,
profile, inversion_policy, periodic_x), the same load / material / collider / step /
export methods and readouts, and the same exceptions. Underneath, ``HipSimulator`` keeps
every array as a torch tensor on the GPU and calls the gfx950 kernels of libmpm_hip.so
through the C ABI (include/mpm_hip.h):

* one tick = a prologue particle pass, S grid updates, S - 1 fused g2p+stress+p2g passes
  and an epilogue gather, enqueued back to back on one stream with no host work between
  substeps: collider poses for the whole tick are precomputed into a device table
  (core.colliders) and the launch box lives in device memory;
* the whole tick can therefore be captured once into a hipGraph and replayed;
* the per-tick guard reads back 8 floats from a device reduction instead of copying
  all positions and velocities to the host.

There is no CPU execution path: ``device="cpu"`` raises, and a missing libmpm_hip.so
raises at load.
"""
from __future__ import annotations

import ctypes as C
import math
import warnings
from dataclasses import dataclass, field
from typing import Any

import numpy as np

from . import native
from .colliders import Collider, ColliderSet, grid_box

f32 = np.float32
MATERIAL_NAME_TO_ID = {
    "jelly": 0, "metal": 1, "sand": 2, "plasticine": 5, "fluid": 6, "stationary": 7,
    "rigid": 8, "mu_i_sand": 9, "newtonian": 10, "mu_i_phi": 11, "tabulated_viscous": 12,
    "tabulated_mu_i": 13,
}
_COV6_IDX = ((0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2))


def _cov_to_six(cov) -> np.ndarray:
    cov = np.asarray(cov)
    if cov.ndim == 2 and cov.shape[1] == 6:
        return np.ascontiguousarray(cov, dtype=np.float32)
    if cov.ndim == 3 and cov.shape[1:] == (3, 3):
        return np.ascontiguousarray(np.stack([cov[:, i, j] for i, j in _COV6_IDX], 1), np.float32)
    raise ValueError(f"cov must be (N, 6) or (N, 3, 3), got shape {cov.shape}")


@dataclass
class GridConfig:
    n_grid: int = 64
    grid_lim: float = 0.4

    @property
    def dx(self) -> float:
        return self.grid_lim / self.n_grid


def _resolve_device(device: str) -> str:
    import torch

    if device in ("cpu",):
        raise NotImplementedError(
            "this backend runs only on the GPU (HIP); device='cpu' has no implementation here")
    if not torch.cuda.is_available():
        raise RuntimeError("no HIP device is visible: the HIP Solver needs a GPU")
    if device in ("auto", "hip", "cuda", "gpu"):
        return "cuda:0"
    if device.startswith("hip:"):
        return "cuda:" + device.split(":", 1)[1]
    if device.startswith("cuda"):
        return device
    raise ValueError(f"unknown device {device!r}")


class HipSimulator:
    """Device state and kernel orchestration (the role of MPM_Simulator_WARP)."""

    def __init__(self, pos, vol, n_grid: int, grid_lim: float, device: str, cov=None,
                 scene=None, n_scenes: int = 1):
        import torch

        self.torch = torch
        self.lib = native.load()
        self.device = torch.device(device)
        n = int(len(pos))
        self.n_particles = n
        self.n_grid = int(n_grid)
        self.dx = float(f32(grid_lim / n_grid))
        self.inv_dx = float(f32(float(n_grid / grid_lim)))
        self.grid_lim = float(grid_lim)
        dev = self.device
        t = lambda *s, dt=torch.float32: torch.zeros(*s, dtype=dt, device=dev)  # noqa: E731
        self.x = torch.from_numpy(np.ascontiguousarray(np.asarray(pos, np.float32).T)).to(dev).contiguous()
        self.v = t(3, n)
        eye = torch.zeros(9, n, dtype=torch.float32, device=dev)
        eye[0] = eye[4] = eye[8] = 1.0
        self.F = eye.clone()
        self.F_trial = eye.clone()
        self.C = t(9, n)
        self.stress = t(9, n)
        self.L = t(9, n)
        self.vol = torch.from_numpy(np.asarray(vol, np.float32).copy()).to(dev)
        self.mass = t(n)
        self.density = t(n)
        self.mu = t(n)
        self.lam = t(n)
        self.yield_stress = t(n)
        self.Jp = t(n)
        self.material = t(n, dt=torch.int32)
        self.selection = t(n, dt=torch.int32)
        # batched scenes: one grid per scene back to back, scene id per particle
        self.n_scenes = int(n_scenes)
        self.scene = None
        if scene is not None:
            self.scene = torch.from_numpy(np.asarray(scene, np.int32).copy()).to(dev)
        elif self.n_scenes > 1:
            raise ValueError("n_scenes > 1 needs a scene id per particle")
        self.init_cov = None
        self.cov = None
        if cov is not None:
            self.init_cov = torch.from_numpy(np.asarray(cov, np.float32).reshape(n, 6).copy()).to(dev)
        self.update_cov_with_F = 0
        # blocked dense grid
        nb = (self.n_grid + 3) // 4
        self.nb = (nb, nb, nb)
        self.blocks_per_scene = nb ** 3
        self.n_blocks = self.blocks_per_scene * self.n_scenes
        self.node_in = t(self.n_blocks * 64 * 4)
        self.node_out = t(self.n_blocks * 64 * 4)
        # per-block scatter / write state: the grid update skips blocks nothing reached
        self.block_touch = t(self.n_blocks, dt=torch.int32)
        self.error_flag = t(1, dt=torch.int32)
        self.bounds = t(9)  # mpm_bounds' 8 floats, then the error flag's bits (one readback)
        # block lists and device tick control
        self.box_list = t(self.n_blocks, dt=torch.int32)
        self.ctl = t(native.TICK_CTL_DTYPE.itemsize // 4, dt=torch.int32)
        self._ctl_host = torch.zeros_like(self.ctl, device="cpu").pin_memory()
        self._h2d_done = None  # event after the last pinned-host upload
        self._blk = None
        self.sparse_ready = False
        # model
        self.model = np.zeros(1, dtype=native.MODEL_DTYPE)
        m = self.model[0]
        m["muI_I0"][:] = 1.0
        m["muI_d"][:] = 1.0
        m["muI_rho_s"][:] = 1.0
        m["muI_phi_init"][:] = 1.0
        m["softening"][:] = 0.1
        sin_phi = f32(np.sin(f32(25.0 / 180.0 * 3.14159265)))
        m["alpha"][:] = f32(np.sqrt(f32(2.0 / 3.0)) * f32(2.0) * sin_phi / (f32(3.0) - sin_phi))
        m["grid_v_damping_scale"] = 1.1
        m["eta_table_smin"], m["eta_table_smax"], m["eta_table_n"] = -1.0, 2.0, 2
        self.eta_table = t(2)
        m["eta_table"] = self.eta_table.data_ptr()
        self.model_dev = torch.zeros(native.MODEL_DTYPE.itemsize, dtype=torch.uint8, device=dev)
        self._model_dirty = True
        self.material_id = 0
        self._mat_ids = {0}  # material ids present (selects the pass specialisation)
        self._m_max = None  # heaviest particle (deterministic-scatter scale), cached
        # colliders
        self.colliders = ColliderSet(self.n_grid, self.dx)
        self.scene_colliders = None  # batched scenes: a colliders.SceneColliders
        self.n_slots = 0
        self.wrench = t(6 * 64)
        self.launch_bits = 0  # particle-pass launch shape (PASS_WG128 / PASS_PPT2 / PASS_PPT4)
        self.grid_ws = None
        self.sched_dev = None
        self._sched_host = None
        self.time = 0.0
        self.grid_launch_box = None
        self._prev_grid_box = None
        self.restrict_grid = True
        self.profile = False
        self.time_profile: dict[str, list[float]] = {}
        self.use_graph = True
        self._graph = None
        self._graph_key = None
        self._ticks = 0
        self.cdf = None      # CdfColliders, created by the first CPIC collider
        self.cdf_tag = None  # per-particle CPIC tag (allocated with it)
        # rigid bodies (material 8): per-particle body id and body-frame reference
        # position, per-body state, restitution planes (mpm_solver_warp.py:366-371, 486-504)
        self.rigid_id = None
        self.x_ref = None
        self.n_rigid_bodies = 0
        self.rigid_bodies = None    # device bytes of mpm_rigid_body_t[n_rigid_bodies]
        self.rigid_planes = []      # host rows of mpm_rigid_plane_t
        self._rigid_planes_dev = None
        self._rigid_planes_key = None
        self.rigid_ws = None
        # particle modifiers (pre-P2G impulses, velocity modifiers), one mask bit each
        self.modifier_mask = None
        self.modifiers = []
        self._modifiers_dev = None
        self._modifiers_key = None
        # device int32 [S, 2] per tick: bit j of column 0 = modifier j active in substep k,
        # of column 1 = restitution plane j active (windows evaluated in float64 on the
        # host, as the reference's host loop compares its times)
        self.substep_active = None
        self._windows = ([], [])  # (modifier, plane) (start, end) in float64
        # deterministic scatter (fixed-point integer accumulation), off by default
        self.deterministic = False
        self.node_fix = None
        self._prepared = None  # prepare_tick's collider rows for the coming tick
        self.chunked = True  # group-aligned chunks at every sort (False: fixed runs of 256)
        self._sort_buf = None  # sort_by_block's work buffers
        self._sort_graph = None  # its captured permutation (key: buffers and flags)
        self._sort_graph_key = None
        # particle-pass chunks (mpm_sort_chunks), computed at every sort; None until then
        self.chunk_start = None
        self.chunk_count = None  # device int32 [2]: chunks of <= 256, of <= 512 particles
        self.chunk_capacity = 0
        # paired interior pass (PASS_PAIR): two particles per lane over chunks of <= 512,
        # for the deterministic lean substep when no launch shape is forced
        self.pair_pass = False
        self.chunk_start2 = None
        self.chunk_capacity2 = 0
        # block pool (enable_pool): node storage for the active blocks only
        self.pool = False
        self.block_slot = None
        self.pool_list = None
        self.pool_blocks = 0
        self._pool_temp = None
        self._vmax = 0.0
        self._fix_exp = None
        self._structs()

    # --------------------------------------------------------------- ABI structs
    def _structs(self):
        P = native.Particles()
        P.n, P.stride = self.n_particles, self.n_particles
        for nm in ("x", "v", "F", "F_trial", "C", "stress", "L", "vol", "mass", "mu", "lam",
                   "yield_stress", "Jp", "material", "selection"):
            setattr(P, nm, native.ptr(getattr(self, nm)))
        P.cov = native.ptr(self.cov)
        P.scene = native.ptr(self.scene)
        P.cdf_tag = native.ptr(self.cdf_tag)
        P.rigid_id = native.ptr(self.rigid_id)
        P.x_ref = native.ptr(self.x_ref)
        P.modifier_mask = native.ptr(self.modifier_mask)
        P.chunk_start = native.ptr(self.chunk_start)
        P.chunk_count = native.ptr(self.chunk_count)
        P.chunk_capacity = self.chunk_capacity
        P2 = native.Particles()  # the paired pass's view: the same arrays, 512-particle chunks
        C.memmove(C.byref(P2), C.byref(P), C.sizeof(P))
        P2.chunk_start = native.ptr(self.chunk_start2)
        P2.chunk_count = (native.ptr(self.chunk_count) + 4) if self.chunk_count is not None else None
        P2.chunk_capacity = self.chunk_capacity2
        G = native.Grid()
        G.node_in, G.node_out = native.ptr(self.node_in), native.ptr(self.node_out)
        G.n[:] = [self.n_grid] * 3
        G.nb[:] = list(self.nb)
        G.dx, G.inv_dx = self.dx, self.inv_dx
        G.error_flag = native.ptr(self.error_flag)
        G.n_scenes = self.n_scenes
        if self.cdf is not None:
            self.cdf.fill(G)
        G.node_fix = native.ptr(self.node_fix)
        G.block_touch = native.ptr(self.block_touch)
        G.block_slot = native.ptr(self.block_slot)
        G.pool_blocks = self.pool_blocks
        if self._fix_exp is not None:
            G.fix_scale[0] = math.ldexp(1.0, self._fix_exp[0])
            G.fix_scale[1] = math.ldexp(1.0, self._fix_exp[1])
        self.P, self.P2, self.G = P, P2, G
        self.pP, self.pP2, self.pG = C.byref(P), C.byref(P2), C.byref(G)

    def stream(self) -> int:
        return self.torch.cuda.current_stream(self.device).cuda_stream

    def _upload_model(self):
        if self._model_dirty:
            host = self.torch.from_numpy(self.model.view(np.uint8).copy())
            self.model_dev.copy_(host, non_blocking=False)
            self._model_dirty = False

    # --------------------------------------------------------------- parameters
    def _material_name_to_id(self, name):
        if isinstance(name, (int, np.integer)):
            return int(name)
        if name not in MATERIAL_NAME_TO_ID:
            raise TypeError(f"Undefined material type: {name}")
        return MATERIAL_NAME_TO_ID[name]

    def _set_type_params(self, m: int, kw: dict):
        md = self.model[0]
        for key, fld in (("E", "E"), ("nu", "nu"), ("bulk_modulus", "bulk"),
                         ("xi", "xi"), ("plastic_viscosity", "plastic_viscosity"),
                         ("softening", "softening"), ("mu_s", "muI_mu_s"),
                         ("delta_mu", "muI_delta_mu"), ("I0", "muI_I0"),
                         ("grain_diameter", "muI_d"), ("grain_density", "muI_rho_s"),
                         ("phi_init", "muI_phi_init"), ("phi_chi", "muI_phi_chi")):
            if key in kw:
                md[fld][m] = float(kw[key])
        if "hardening" in kw:
            md["hardening"][m] = float(kw["hardening"])
        if "friction_angle" in kw:
            sin_phi = float(np.sin(kw["friction_angle"] / 180.0 * 3.14159265))
            md["alpha"][m] = np.sqrt(2.0 / 3.0) * 2.0 * sin_phi / (3.0 - sin_phi)
        self._model_dirty = True

    def set_parameters_dict(self, kw: dict):
        """set_parameters_dict (mpm_solver_warp.py:634-800)."""
        torch = self.torch
        if "material" in kw:
            self.material_id = self._material_name_to_id(kw["material"])
            self.material.fill_(self.material_id)
            self._mat_ids = {self.material_id}
        # the reference reallocates (zeroes) the grid on every call
        self.node_in.zero_()
        self.node_out.zero_()
        m = self.material_id
        self._set_type_params(m, kw)
        md = self.model[0]
        if "yield_stress" in kw:
            self.yield_stress.fill_(float(kw["yield_stress"]))
        if "eta_table" in kw:
            tab = np.asarray(kw["eta_table"], dtype=np.float32).reshape(-1)
            self.eta_table = torch.from_numpy(tab.copy()).to(self.device)
            md["eta_table"] = self.eta_table.data_ptr()
            md["eta_table_n"] = int(tab.shape[0])
        if "eta_table_smin" in kw:
            md["eta_table_smin"] = float(kw["eta_table_smin"])
        if "eta_table_smax" in kw:
            md["eta_table_smax"] = float(kw["eta_table_smax"])
        if "g" in kw:
            md["gravity"] = np.asarray(kw["g"], dtype=np.float64).astype(f32)
        if "density" in kw:
            self.density.fill_(float(kw["density"]))
            torch.mul(self.density, self.vol, out=self.mass)
            self._m_max = None  # the fixed-point mass scale follows the heaviest particle
        if "E" in kw or "nu" in kw:
            self.finalize_mu_lam()
        if "rpic_damping" in kw:
            md["rpic_damping"] = kw["rpic_damping"]
        if "grid_v_damping_scale" in kw:
            md["grid_v_damping_scale"] = kw["grid_v_damping_scale"]
        self._model_dirty = True

    def finalize_mu_lam(self):
        """compute_mu_lam_from_E_nu (mpm_utils.py:583-588), fp32 on device."""
        self._m_max = None
        torch = self.torch
        md = self.model[0]
        E = torch.from_numpy(md["E"].copy()).to(self.device)[self.material.long()]
        nu = torch.from_numpy(md["nu"].copy()).to(self.device)[self.material.long()]
        self.mu.copy_(E / (2.0 * (1.0 + nu)))
        self.lam.copy_(E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu)))

    def set_parameters_for_particles(self, a: int, b: int, kw: dict):
        """set_parameters_for_particles (mpm_solver_warp.py:1700-1783)."""
        self._m_max = None
        m = None
        if "material" in kw:
            m = self._material_name_to_id(kw["material"])
            self.material[a:b] = m
            self._mat_ids.add(m)
        if m is None:
            m = self.material_id
        self._set_type_params(m, kw)
        if "density" in kw:
            self.density[a:b] = float(kw["density"])
            self.mass[a:b] = self.density[a:b] * self.vol[a:b]
        if "yield_stress" in kw:
            self.yield_stress[a:b] = float(kw["yield_stress"])
        if "E" in kw or "nu" in kw:
            md = self.model[0]
            E, nu = float(md["E"][m]), float(md["nu"][m])
            self.mu[a:b] = E / (2.0 * (1.0 + nu))
            self.lam[a:b] = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu))
        if "obj_id" in kw:  # rigid body id (mpm_solver_warp.py:1755-1758)
            if self.rigid_id is None:
                self.rigid_id = self.torch.zeros(self.n_particles, dtype=self.torch.int32, device=self.device)
                self._structs()
            self.rigid_id[a:b] = int(kw["obj_id"])

    # --------------------------------------------------------------- rigid bodies
    def finalize_rigid_bodies(self):
        """finalize_rigid_bodies (mpm_solver_warp.py:834-873): per body (material 8,
        grouped by obj_id) mass, centre of mass (float32), inverse body-frame inertia
        (summed in float64) and each particle's body-frame reference position;
        orientation I, at rest."""
        torch = self.torch
        mat = self.material.cpu().numpy()
        rigid = mat == 8
        if not np.any(rigid):
            return
        if self.rigid_id is None:
            self.rigid_id = torch.zeros(self.n_particles, dtype=torch.int32, device=self.device)
        rid = self.rigid_id.cpu().numpy()
        x = np.ascontiguousarray(self.x.t().cpu().numpy())
        m = self.mass.cpu().numpy()
        nb = int(rid[rigid].max()) + 1
        body = np.zeros(nb, dtype=native.RIGID_BODY_DTYPE)
        x_ref = np.zeros_like(x)
        for b in range(nb):
            idx = np.where(rigid & (rid == b))[0]
            if len(idx) == 0:
                raise ValueError(f"rigid body {b} has no particles (obj_id values must be 0..{nb - 1})")
            m_b = m[idx]
            mass = f32(m_b.sum())
            x_cm = ((x[idx] * m_b[:, None]).sum(axis=0) / mass).astype(f32)
            r = x[idx] - x_cm
            x_ref[idx] = r
            # m (r.r I - r r^T): float32 dot and outer, float64 eye and sum, as the
            # reference's loop evaluates it
            rr = np.einsum("ij,ij->i", r, r).astype(np.float64)
            outer = (r[:, :, None] * r[:, None, :]).astype(np.float64)
            inertia = (m_b.astype(np.float64)[:, None, None] * (rr[:, None, None] * np.eye(3) - outer)).sum(axis=0)
            body[b]["x_cm"] = x_cm
            body[b]["mass"] = mass
            body[b]["R"] = np.eye(3, dtype=f32).reshape(-1)
            body[b]["inv_inertia"] = np.linalg.inv(inertia).astype(f32).reshape(-1)
        self.rigid_bodies = torch.from_numpy(body.view(np.uint8).copy()).to(self.device)
        self.x_ref = torch.from_numpy(np.ascontiguousarray(x_ref.T)).to(self.device)
        self.n_rigid_bodies = nb
        self._rigid_ws_ensure()
        self._structs()
        self._graph = None

    def _rigid_ws_ensure(self):
        nbytes = int(self.lib.mpm_rigid_workspace_bytes(self.n_rigid_bodies, len(self.rigid_planes)))
        if self.rigid_ws is None or self.rigid_ws.numel() < max(nbytes, 8):
            self.rigid_ws = self.torch.zeros(max(nbytes, 8), dtype=self.torch.uint8, device=self.device)
            self._graph = None

    def add_rigid_plane(self, point, normal, restitution, friction, start_time, end_time):
        """A restitution surface (mpm_solver_warp.py:1915-1925); `normal` already unit."""
        row = np.zeros(1, dtype=native.RIGID_PLANE_DTYPE)
        row["point"], row["normal"] = point, normal
        row["restitution"], row["friction"] = restitution, friction
        row["start_time"], row["end_time"] = start_time, min(end_time, 3.0e38)
        if len(self.rigid_planes) >= 32:
            raise ValueError("at most 32 restitution planes")
        self.rigid_planes.append(row)
        self._windows[1].append((float(start_time), float(end_time)))
        if self.n_rigid_bodies:
            self._rigid_ws_ensure()

    def _rigid_planes_device(self):
        if not self.rigid_planes:
            return None
        key = len(self.rigid_planes)
        if self._rigid_planes_key != key:
            rows = np.concatenate(self.rigid_planes)
            self._rigid_planes_dev = self.torch.from_numpy(rows.view(np.uint8).copy()).to(self.device)
            self._rigid_planes_key = key
            self._graph = None
        return self._rigid_planes_dev

    def rigid_state(self, body: int = 0) -> dict:
        if not 0 <= body < self.n_rigid_bodies:
            raise ValueError(f"no rigid body {body} (finalize_rigid_bodies found {self.n_rigid_bodies})")
        b = self.rigid_bodies.cpu().numpy().view(native.RIGID_BODY_DTYPE)[body]
        return {"com": b["x_cm"].copy(), "v": b["v_cm"].copy(), "omega": b["omega"].copy(),
                "R": b["R"].reshape(3, 3).copy()}

    def set_rigid_body_velocity(self, body_id, v_cm=(0.0, 0.0, 0.0), omega=(0.0, 0.0, 0.0)):
        """set_rigid_body_velocity (mpm_solver_warp.py:874-880)."""
        host = self.rigid_bodies.cpu().numpy().copy()
        rows = host.view(native.RIGID_BODY_DTYPE)
        rows[body_id]["v_cm"] = v_cm
        rows[body_id]["omega"] = omega
        self.rigid_bodies.copy_(self.torch.from_numpy(host))

    def _rigid(self, k, dt, s):
        planes = self._rigid_planes_device()
        native.check(self.lib.mpm_rigid_step(
            self.pP, self.pG, native.ptr(self.rigid_bodies), self.n_rigid_bodies, native.ptr(planes),
            len(self.rigid_planes), native.ptr(self.substep_active) + 8 * k + 4, dt, native.ptr(self.rigid_ws), s),
            "rigid_step")

    # --------------------------------------------------------------- particle modifiers
    def _add_modifier(self, row, t0: float, t1: float):
        if len(self.modifiers) >= native.MAX_MODIFIERS:
            raise ValueError(f"at most {native.MAX_MODIFIERS} particle modifiers")
        if self.modifier_mask is None:
            self.modifier_mask = self.torch.zeros(self.n_particles, dtype=self.torch.int32, device=self.device)
            self._structs()
        row["bit"] = len(self.modifiers)
        native.check(self.lib.mpm_modifier_select(self.pP, row.ctypes.data, self.stream()), "modifier_select")
        self.modifiers.append(row)
        self._windows[0].append((float(t0), float(t1)))
        self._graph = None

    def add_impulse_on_particles(self, force, dt, point=(1, 1, 1), size=(1, 1, 1), num_dt=1, start_time=0.0,
                                 device=None):
        """add_impulse_on_particles (mpm_solver_warp.py:2385-2427): particles in the box
        |x - point| < size (selected now) get v += force / mass * dt while
        start_time <= t < start_time + dt * num_dt."""
        row = np.zeros(1, dtype=native.MODIFIER_DTYPE)
        row["kind"] = 0
        row["start_time"], row["end_time"] = start_time, start_time + dt * num_dt
        row["point"], row["size"], row["vec"] = point, size, force
        self._add_modifier(row, start_time, start_time + dt * num_dt)

    def enforce_particle_velocity_translation(self, point, size, velocity, start_time, end_time, device=None):
        """enforce_particle_velocity_translation (mpm_solver_warp.py:2431-2469): particles in
        the box (selected now) move at `velocity` while start_time <= t < end_time."""
        row = np.zeros(1, dtype=native.MODIFIER_DTYPE)
        row["kind"] = 1
        row["start_time"], row["end_time"] = start_time, end_time
        row["point"], row["size"], row["vec"] = point, size, velocity
        self._add_modifier(row, start_time, end_time)

    def enforce_particle_velocity_rotation(self, point, normal, half_height_and_radius, rotation_scale,
                                           translation_scale, start_time, end_time, device=None):
        """enforce_particle_velocity_rotation (mpm_solver_warp.py:2475-2539): particles in
        the cylinder about `normal` through `point` (selected now) rotate about the axis at
        rotation_scale rad/s and translate along it at translation_scale."""
        n = np.asarray(normal, np.float64)
        n = (1.0 / math.sqrt(float((n * n).sum()))) * n
        nf = n.astype(f32)
        h1 = np.ones(3, f32)
        if abs(float(np.dot(nf, h1))) < 0.01:
            h1 = np.array([0.72, 0.37, -0.67], f32)
        h1 = (h1 - np.dot(h1, nf) * nf).astype(f32)
        h1 = (h1 * (f32(1.0) / np.sqrt(np.dot(h1, h1)))).astype(f32)
        h2 = np.cross(h1, nf).astype(f32)
        row = np.zeros(1, dtype=native.MODIFIER_DTYPE)
        row["kind"] = 2
        row["start_time"], row["end_time"] = start_time, end_time
        row["point"], row["vec"], row["axis1"], row["axis2"] = point, nf, h1, h2
        row["half_height"], row["radius"] = half_height_and_radius[0], half_height_and_radius[1]
        row["rotation_scale"], row["translation_scale"] = rotation_scale, translation_scale
        self._add_modifier(row, start_time, end_time)

    def _modifiers_device(self):
        key = len(self.modifiers)
        if self._modifiers_key != key:
            rows = np.concatenate(self.modifiers)
            self._modifiers_dev = self.torch.from_numpy(rows.view(np.uint8).copy()).to(self.device)
            self._modifiers_key = key
            self._graph = None
        return self._modifiers_dev

    def _apply_modifiers(self, k, s):
        native.check(self.lib.mpm_modifiers_apply(
            self.pP, native.ptr(self._modifiers_device()), len(self.modifiers),
            native.ptr(self.substep_active) + 8 * k, self._dt, s), "modifiers_apply")

    def set_gravity(self, g):
        self.model[0]["gravity"] = np.asarray(g, dtype=np.float64).astype(f32)
        self._model_dirty = True

    def enable_cov_step(self):
        self.update_cov_with_F = 1
        self.model[0]["update_cov_with_F"] = 1
        self._model_dirty = True
        self.cov = self.init_cov.t().contiguous()
        self._structs()
        self._graph = None

    # --------------------------------------------------------------- grid boxes
    def _grid_box(self, lo_w, hi_w, halo=1):
        return grid_box(lo_w, hi_w, self.dx, (self.n_grid,) * 3, halo)

    def _boxes(self):
        """Current particle box and the zero box (union with the previous tick's),
        mpm_solver_warp.py:1415-1432."""
        full = ((0, 0, 0), (self.n_grid,) * 3)
        gbox = self.grid_launch_box if (self.restrict_grid and self.grid_launch_box) else None
        if gbox is None:
            self._prev_grid_box = None
            return full, full
        prev = self._prev_grid_box
        if prev is None:
            z = gbox
        else:
            zl = tuple(min(prev[0][i], gbox[0][i]) for i in range(3))
            zh = tuple(max(prev[0][i] + prev[1][i], gbox[0][i] + gbox[1][i]) for i in range(3))
            z = (zl, tuple(zh[i] - zl[i] for i in range(3)))
        self._prev_grid_box = gbox
        return gbox, z

    # --------------------------------------------------------------- guard
    def read_bounds(self):
        """(min xyz, max xyz, max |v|, non-finite count, error flag) via a device
        reduction; one small device->host copy."""
        native.check(self.lib.mpm_bounds(self.pP, native.ptr(self.bounds), self.stream()), "bounds")
        self.bounds[8:9].copy_(self.error_flag.view(self.torch.float32))
        b = self.bounds.cpu().numpy()
        flag = int(b[8:9].view(np.int32)[0])
        self._vmax = float(b[6])
        return b[0:3].copy(), b[3:6].copy(), float(b[6]), int(b[7]), flag

    # --------------------------------------------------------------- deterministic mode
    def enable_deterministic(self):
        """Switch the scatter to 64-bit fixed-point integer accumulation (bit-reproducible
        grid sums, independent of atomic order, launch shape and particle order)."""
        if self.cdf is not None and self.cdf.n > 0:
            raise NotImplementedError("the deterministic scatter does not cover CDF colliders")
        self.deterministic = True
        if self.node_fix is None:
            blocks = self.pool_blocks if self.pool else self.n_blocks
            self.node_fix = self.torch.zeros(blocks * 64 * 4, dtype=self.torch.int64, device=self.device)
        self.grid_ws = None  # the fixed-point wrench needs a zeroed workspace
        self._graph = None
        self._structs()

    def disable_deterministic(self):
        """Back to float-atomic accumulation of the scatter and the wrench."""
        self.deterministic = False
        self.grid_ws = None
        self._graph = None
        self._structs()

    def _fix_scales(self, dt: float):
        """Powers of two such that the largest expected contribution is about 2^40 units
        (2^11 below the per-contribution limit): mass by the heaviest particle, momentum by
        it times 2 dx/dt (a stable explicit substep moves nothing a cell; the scale then
        depends on dt and the masses only, so captured graphs stay valid tick to tick)."""
        if self._m_max is None:
            self._m_max = float(self.mass.max().item()) if self.n_particles else 1.0
        m = max(self._m_max, 1e-30)
        p = m * 2.0 * self.dx / dt
        exp = (40 - math.ceil(math.log2(m)), 40 - math.ceil(math.log2(p)))
        if exp != self._fix_exp:
            self._fix_exp = exp
            self._structs()

    # --------------------------------------------------------------- block pool
    def enable_pool(self, initial_blocks: int = 64):
        """Node storage for the active blocks only (mpm_pool_assign): node_in, node_out and
        node_fix hold `pool_blocks` blocks, handed out again at every active-block rebuild
        and grown (never shrunk) when the active set outgrows them. Needs the sparse
        stepping (a rebuild before every tick)."""
        if self.cdf is not None or self.n_rigid_bodies:
            raise NotImplementedError("the block pool does not cover CPIC colliders or rigid bodies")
        torch = self.torch
        self.pool = True
        self.block_slot = torch.full((self.n_blocks,), -1, dtype=torch.int32, device=self.device)
        self.pool_list = torch.zeros(self.n_blocks, dtype=torch.int32, device=self.device)
        tb = int(self.lib.mpm_pool_temp_bytes(self.n_blocks))
        if tb < 0:
            raise RuntimeError("mpm_pool_temp_bytes failed")
        self._pool_temp = torch.empty(max(tb, 1), dtype=torch.uint8, device=self.device)
        self._pool_resize(initial_blocks)
        self.sparse_ready = False

    def _pool_resize(self, blocks: int):
        torch = self.torch
        blocks = max(64, (int(blocks) + 63) // 64 * 64)
        self.pool_blocks = blocks
        self.node_in = torch.zeros(blocks * 64 * 4, dtype=torch.float32, device=self.device)
        self.node_out = torch.zeros(blocks * 64 * 4, dtype=torch.float32, device=self.device)
        if self.node_fix is not None:
            self.node_fix = torch.zeros(blocks * 64 * 4, dtype=torch.int64, device=self.device)
        self._graph = None
        self._structs()

    def grid_bytes(self) -> int:
        """Device bytes of the grid storage: node arrays, block states and the pool table."""
        ts = [self.node_in, self.node_out, self.node_fix, self.block_touch, self.block_slot, self.pool_list]
        if self._blk is not None:
            ts += list(self._blk.values())
        return int(sum(t.numel() * t.element_size() for t in ts if t is not None))

    # --------------------------------------------------------------- sparse blocks
    def rebuild_active_blocks(self):
        """rebuild_active_blocks (mpm_solver_warp.py:989-1026), fully on device: the
        counts stay on the GPU and size nothing on the host."""
        torch = self.torch
        if self._blk is None:
            nb = self.n_blocks
            z = lambda: torch.zeros(nb, dtype=torch.int32, device=self.device)  # noqa: E731
            self._blk = {"raw": z(), "cur": z(), "prev": z(), "cur_list": z(),
                         "union_list": z(),
                         "counts": torch.zeros(2, dtype=torch.int32, device=self.device)}
        b = self._blk
        b["prev"], b["cur"] = b["cur"], b["prev"]
        b["raw"].zero_()
        b["counts"].zero_()
        s = self.stream()
        L = self.lib
        native.check(L.mpm_blocks_mark(self.pP, self.pG, native.ptr(b["raw"]), s), "blocks_mark")
        native.check(L.mpm_blocks_dilate(self.pG, native.ptr(b["raw"]), native.ptr(b["cur"]), s),
                     "blocks_dilate")
        if self.pool:
            self._pool_assign(b, s)
            return
        native.check(L.mpm_blocks_compact(self.pG, native.ptr(b["cur"]), native.ptr(b["prev"]),
                                          native.ptr(b["cur_list"]), native.ptr(b["union_list"]),
                                          native.ptr(b["counts"]), s), "blocks_compact")
        self.sparse_ready = True

    def _pool_assign(self, b, s):
        """Slots for this tick's active blocks (in block-code order; the grid update walks
        them as its list), the node arrays grown to hold them, and the slots in use
        cleared: a slot may have held another block last tick."""
        native.check(self.lib.mpm_pool_assign(self.pG, native.ptr(b["cur"]), native.ptr(self.block_slot),
                                              native.ptr(self.pool_list), native.ptr(b["counts"]),
                                              native.ptr(self._pool_temp), self._pool_temp.numel(), s),
                     "pool_assign")
        count = int(b["counts"][0].item())
        if count > self.pool_blocks:
            self._pool_resize(count + count // 4)
        used = count * 64 * 4
        self.node_in[:used].zero_()
        self.node_out[:used].zero_()
        if self.node_fix is not None:
            self.node_fix[:used].zero_()
        self.block_touch.zero_()
        self.sparse_ready = True

    def active_block_fraction(self) -> float:
        if not self.sparse_ready:
            return 1.0
        return int(self._blk["counts"][0].item()) / float(self.n_blocks)

    # --------------------------------------------------------------- sort
    def sort_by_block(self, cells: bool = False) -> bool:
        """Stable block-key sort on device (solver.py:454-467, mpm_solver_warp.py:
        1548-1619); returns False (no permutation) when already ordered, read back from a
        device flag right after the key kernel. With `cells` the key is refined by the
        stencil base cell inside the block. The work buffers are allocated once, and the
        work after the flag (radix sort, permutation, chunk boundaries: ~80 launches) is
        replayed as one captured hipGraph."""
        torch = self.torch
        n = self.n_particles
        if n < 2:
            return False
        dev = self.device
        if self._sort_buf is None:
            tb = int(self.lib.mpm_sort_temp_bytes(n))
            i32 = dict(dtype=torch.int32, device=dev)
            self._sort_buf = {"keys": torch.empty(n, **i32), "vals": torch.empty(n, **i32),
                              "flag": torch.zeros(1, **i32), "keys_o": torch.empty(n, **i32),
                              "perm": torch.empty(n, **i32), "tb": tb,
                              "perm_l": torch.empty(n, dtype=torch.int64, device=dev),
                              "temp": torch.empty(max(tb, 1), dtype=torch.uint8, device=dev),
                              "scratch": torch.empty(9 * n, dtype=torch.float32, device=dev),
                              "density": torch.empty_like(self.density)}
        b = self._sort_buf
        if self.chunked:
            self._chunk_buffers()
        s = self.stream()
        keyfn = self.lib.mpm_sort_keys_cells if cells else self.lib.mpm_sort_keys
        b["flag"].zero_()
        native.check(keyfn(self.pP, self.pG, native.ptr(b["keys"]), native.ptr(b["vals"]), native.ptr(b["flag"]), s),
                     "sort_keys")
        if int(b["flag"].item()) == 0:
            if self.chunked:
                self._chunk_kernels(b["keys"], cells, s)
                self._chunk_counts()
            return False
        if self.use_graph:
            key = (bool(cells), n, self.chunked, C.string_at(C.addressof(self.P), C.sizeof(self.P)),
                   native.ptr(self.density), native.ptr(self.init_cov))
            if self._sort_graph is None or self._sort_graph_key != key:
                g = torch.cuda.CUDAGraph()
                torch.cuda.synchronize(dev)
                with torch.cuda.graph(g):
                    self._permute_sorted(cells, self.stream())
                self._sort_graph, self._sort_graph_key = g, key
            self._sort_graph.replay()
        else:
            self._permute_sorted(cells, s)
        self._last_perm = b["perm_l"]  # valid until the next sort
        if self.chunked:
            self._chunk_counts()
        return True

    def _permute_sorted(self, cells: bool, s):
        """Radix sort of the keys, every particle array permuted in place, the host-side
        per-particle arrays alongside, then the chunk boundaries of the sorted order."""
        b = self._sort_buf
        n = self.n_particles
        native.check(self.lib.mpm_sort_pairs(native.ptr(b["keys"]), native.ptr(b["keys_o"]), native.ptr(b["vals"]),
                                             native.ptr(b["perm"]), n, native.ptr(b["temp"]), b["tb"], s),
                     "sort_pairs")
        native.check(self.lib.mpm_permute(self.pP, native.ptr(b["perm"]), native.ptr(b["scratch"]), s), "permute")
        pl = b["perm_l"]
        pl.copy_(b["perm"])
        self.torch.index_select(self.density, 0, pl, out=b["density"])
        self.density.copy_(b["density"])
        if self.init_cov is not None:
            self.init_cov.copy_(self.init_cov[pl])
        if self.chunked:
            self._chunk_kernels(b["keys_o"], cells, s)

    def sync_chunk_capacity(self):
        """Size the particle-pass launches from the current chunk counts (waits for them):
        for callers that sort once and then run many substeps on that order."""
        if self.chunk_start is not None:
            self._chunk_counts(wait=True)

    def _chunk_buffers(self):
        torch = self.torch
        n = self.n_particles
        if self.chunk_start is None:
            self.chunk_start = torch.zeros(n + 1, dtype=torch.int32, device=self.device)
            self.chunk_start2 = torch.zeros(n + 1, dtype=torch.int32, device=self.device)
            self.chunk_count = torch.zeros(2, dtype=torch.int32, device=self.device)
            self._chunk_temp_bytes = int(self.lib.mpm_sort_chunks_temp_bytes(n))
            self._chunk_temp = torch.empty(max(self._chunk_temp_bytes, 1), dtype=torch.uint8, device=self.device)
            self._count_host = torch.zeros(2, dtype=torch.int32).pin_memory()
            self._count_ev = None

    def _chunk_kernels(self, sorted_keys, cells: bool, s):
        """Particle-pass chunks over the sorted particles (mpm_sort_chunks): runs of at most
        256 particles within 4 consecutive blocks of one block column, and runs of at most
        512 for the paired pass."""
        nb = self.n_grid // 4 + 2
        n = self.n_particles
        for size, starts, cnt in ((256, self.chunk_start, native.ptr(self.chunk_count)),
                                  (512, self.chunk_start2, native.ptr(self.chunk_count) + 4)):
            native.check(self.lib.mpm_sort_chunks(native.ptr(sorted_keys), n, 6 if cells else 0, nb, 4, size,
                                                  native.ptr(starts), cnt, native.ptr(self._chunk_temp),
                                                  self._chunk_temp_bytes, s), "sort_chunks")

    def _chunk_counts(self, wait: bool = False):
        """Launch sizes from the chunk counts of an earlier sort, copied back without
        waiting (the pass sweeps its launch again for any chunks beyond it, which costs a
        whole second chain of the pass, so the launch keeps a quarter of the count plus 64
        chunks to spare; surplus workgroups exit at once). The first sort, and `wait`,
        wait for the current counts."""
        torch = self.torch
        if self._count_ev is None or wait:  # wait for the counts
            counts = self.chunk_count.tolist()
        elif self._count_ev.query():  # the previous sort's counts have arrived
            counts = self._count_host.tolist()
        else:
            counts = None
        self._count_host.copy_(self.chunk_count, non_blocking=True)
        self._count_ev = torch.cuda.Event()
        self._count_ev.record()
        if counts is None:
            return
        # launch capacities in steps of 64 chunks, kept while the count stays within them
        # and they hold no more than twice the spare (a new capacity means a new captured
        # graph)
        caps = []
        for count, cap in zip(counts, (self.chunk_capacity, self.chunk_capacity2)):
            spare = count // 4 + 64
            if count > cap or cap - count > 2 * spare + 64:
                cap = (count + spare + 63) // 64 * 64
            caps.append(cap)
        if caps != [self.chunk_capacity, self.chunk_capacity2]:
            self.chunk_capacity, self.chunk_capacity2 = caps
            self._graph = None
            self._structs()

    # --------------------------------------------------------------- ticks
    def _ensure_sched(self, nbytes: int):
        torch = self.torch
        if self.sched_dev is None or self.sched_dev.numel() < nbytes:
            self.sched_dev = torch.zeros(max(nbytes, 256), dtype=torch.uint8, device=self.device)
            self._sched_host = torch.zeros(max(nbytes, 256), dtype=torch.uint8).pin_memory()
            self._graph = None

    def cdf_colliders(self):
        """The CPIC collider set, created (with the per-particle tags) on first use."""
        if self.cdf is None:
            from .cdf import CdfColliders
            self.cdf = CdfColliders(self)
            self.cdf_tag = self.torch.zeros(self.n_particles, dtype=self.torch.int32, device=self.device)
            self._structs()
            self._graph = None
        return self.cdf

    def prepare_tick(self, dt: float, substeps: int):
        """Build the coming tick's collider rows now: host work that, done before the
        guard's readback, overlaps the GPU still running the previous tick. run_tick uses
        them when it is called for the same start time, dt and substep count; otherwise (or
        on drop_prepared) the collider poses the schedule advanced are restored."""
        self.drop_prepared()
        S = int(substeps)
        if S <= 0:
            return
        cs = self.scene_colliders if self.scene_colliders is not None else self.colliders
        snap = cs.poses()
        self._prepared = (self.time, float(dt), S, self._schedule(dt, S), cs, snap)

    def drop_prepared(self):
        p, self._prepared = self._prepared, None
        if p is not None:
            p[4].set_poses(p[5])

    def run_tick(self, dt: float, substeps: int, fused: bool, sparse: bool):
        torch = self.torch
        S = int(substeps)
        if S <= 0:
            self.drop_prepared()
            return
        self._upload_model()
        L = self.lib
        s = self.stream()
        if self.deterministic:
            self._fix_scales(dt)
        self._dt = float(dt)
        if self.modifiers:
            self._modifiers_device()
        if self.n_rigid_bodies:
            self._rigid_planes_device()
            self._rigid_ws_ensure()
        if self.n_rigid_bodies or self.modifiers:
            # rigid bodies and particle modifiers run the split pipeline (solver.py:
            # 441-444); their activity windows against each substep's start time
            fused = False
            if self.substep_active is None or self.substep_active.shape[0] < S:
                self.substep_active = torch.zeros(max(S, 1), 2, dtype=torch.int32, device=self.device)
                self._graph = None
            times = ColliderSet.times(self.time, dt, S)[0]
            act = np.zeros((S, 2), dtype=np.uint32)
            for col, wins in enumerate(self._windows):
                for j, (t0, t1) in enumerate(wins):
                    act[:, col] |= (((times >= t0) & (times < t1)).astype(np.uint32) << np.uint32(j))
            self.substep_active[:S].copy_(torch.from_numpy(act.view(np.int32)))
        p, self._prepared = self._prepared, None
        if p is not None and p[:3] == (self.time, float(dt), S):
            tab, t_end, K, n_slots = p[3]
        else:
            if p is not None:
                p[4].set_poses(p[5])
            tab, t_end, K, n_slots = self._schedule(dt, S)
        self.n_slots = n_slots
        if 6 * n_slots > self.wrench.numel():  # grow, keeping the accumulated values
            w = torch.zeros(6 * n_slots, dtype=torch.float32, device=self.device)
            w[:self.wrench.numel()] = self.wrench
            self.wrench = w
            self._graph = None
        nbytes = tab.nbytes
        self._ensure_sched(nbytes)
        ws = int(L.mpm_grid_workspace_floats(self.n_blocks, n_slots))
        if self.grid_ws is None or self.grid_ws.numel() < max(ws, 1):
            self.grid_ws = torch.zeros(max(ws, 1), dtype=torch.float32, device=self.device)
            self._graph = None
        if self._h2d_done is not None:  # the pinned staging buffers are reused
            self._h2d_done.synchronize()
        if nbytes:
            self._sched_host[:nbytes].numpy()[:] = tab.view(np.uint8).reshape(-1)
            self.sched_dev[:nbytes].copy_(self._sched_host[:nbytes], non_blocking=True)
        ctl = np.zeros(1, dtype=native.TICK_CTL_DTYPE)
        if sparse:
            ctl["box_lo"], ctl["box_hi"] = (0, 0, 0), (self.n_grid,) * 3
            b = self._blk
            zl, zc = b["union_list"], b["counts"][1:]
            glist, gcount = (self.pool_list if self.pool else b["cur_list"]), b["counts"][0:]
            self._prev_grid_box = None
        elif self.pool:
            raise RuntimeError("the block pool needs the sparse stepping (a block rebuild before every tick)")
        else:
            g, z = self._boxes()
            ctl["box_lo"] = g[0]
            ctl["box_hi"] = [g[0][i] + g[1][i] for i in range(3)]
        self._ctl_host.numpy()[:] = ctl.view(np.int32).reshape(-1)
        self.ctl.copy_(self._ctl_host, non_blocking=True)
        self._h2d_done = torch.cuda.Event()
        self._h2d_done.record()
        # clear the zero box (union of this tick's and the previous tick's box)
        if sparse:
            if not self.pool:  # (the pool clears its slots at every rebuild)
                native.check(L.mpm_grid_zero_blocks(self.pG, native.ptr(zl), native.ptr(zc),
                                                    self.n_blocks, 3, s), "zero_blocks")
        else:
            lo = (C.c_int32 * 3)(*z[0])
            hi = (C.c_int32 * 3)(*[z[0][i] + z[1][i] for i in range(3)])
            native.check(L.mpm_grid_zero_box(self.pG, lo, hi, 3, s), "zero_box")
            native.check(L.mpm_blocks_from_box(self.pG, native.ptr(self.ctl), native.ptr(self.box_list), s),
                         "blocks_from_box")
            glist, gcount = self.box_list, self.ctl[6:]
        lists = (native.ptr(glist), native.ptr(gcount))
        cdf = self.cdf if (self.cdf is not None and self.cdf.n > 0) else None
        if cdf is not None:
            self.G.cdf.n_cdf = cdf.n
            self.G.cdf.mask_off = int(cdf.mask_off)
        if cdf is not None and not cdf.quiet(ColliderSet.times(self.time, dt, S)[0]):
            # a CPIC surface moves, restamps or opens / closes this tick: the stamps and
            # lane uploads interleave with the substeps, so the tick runs eagerly
            self._enqueue_tick_cdf(dt, S, K, fused, lists, s)
        elif self.profile:
            self._enqueue_tick_profiled(dt, S, K, fused, lists)
        elif self.use_graph and self._ticks > 0:
            key = (S, float(dt), bool(fused), K, lists, native.ptr(self.sched_dev),
                   self.update_cov_with_F, n_slots, native.ptr(self.grid_ws), native.ptr(self.wrench),
                   self.chunk_capacity, self.chunk_capacity2, self.pair_pass,
                   frozenset(self._mat_ids) <= native.VISCOUS_IDS, self.launch_bits,
                   None if cdf is None else (cdf.n, int(cdf.mask_off), native.ptr(self.cdf_tag)),
                   self._fix_exp if self.deterministic else None,
                   (self.n_rigid_bodies, native.ptr(self.rigid_bodies), native.ptr(self._rigid_planes_dev),
                    len(self.rigid_planes), native.ptr(self.substep_active), native.ptr(self.rigid_ws),
                    native.ptr(self.x_ref)) if self.n_rigid_bodies else None,
                   (len(self.modifiers), native.ptr(self._modifiers_dev), native.ptr(self.modifier_mask),
                    native.ptr(self.substep_active)) if self.modifiers else None)
            if self._graph is None or self._graph_key != key:
                self._capture(dt, S, K, fused, lists)
                self._graph_key = key
            self._graph.replay()
        else:
            self._enqueue_tick(dt, S, K, fused, lists, s)
        self._ticks += 1
        self.time = t_end

    def _capture(self, dt, S, K, fused, lists):
        torch = self.torch
        g = torch.cuda.CUDAGraph()
        torch.cuda.synchronize(self.device)
        with torch.cuda.graph(g):
            self._enqueue_tick(dt, S, K, fused, lists, self.stream())
        self._graph = g

    def _schedule(self, dt, S):
        """Collider rows of one tick: (table [S, n_scenes * K], end time, K colliders per
        scene, wrench slots)."""
        if self.scene_colliders is None:
            tab, t_end = self.colliders.schedule(self.time, dt, S)
            return tab, t_end, tab.shape[1], self.colliders.n_slots
        return self.scene_colliders.schedule(self.time, dt, S)

    def _pass(self, mode, dt, s):
        P = self.pP
        fused = native.PASS_G2P | native.PASS_STRESS | native.PASS_P2G
        if (self.pair_pass and self.deterministic and not self.launch_bits and self.chunk_start2 is not None
                and (mode & native.PASS_LEAN) and (mode & fused) == fused
                and (self.cdf is None or self.cdf.n == 0)):
            P = self.pP2
            mode |= native.PASS_PAIR
        mode |= self.launch_bits
        if self.deterministic:
            mode |= native.PASS_DETERMINISTIC
        if self._mat_ids <= native.VISCOUS_IDS:
            mode |= native.PASS_VISCOUS_ONLY
        native.check(self.lib.mpm_particle_pass(P, self.pG, native.ptr(self.model_dev), dt, mode, s),
                     "particle_pass")

    def _grid(self, dt, k_row, K, lists, s):
        sched = native.ptr(self.sched_dev) + k_row * self.n_scenes * K * native.COLLIDER_DTYPE.itemsize
        slots = self.n_slots
        # deterministic: the integer wrench rows accumulate over the tick and are flushed
        # once at its end (_wrench_flush)
        flags = native.GRID_FIXED_IN | native.GRID_WRENCH_DEFER if self.deterministic else native.GRID_ZERO_IN
        native.check(self.lib.mpm_grid_update(
            self.pG, native.ptr(self.model_dev), native.ptr(self.ctl), lists[0], lists[1],
            self.n_blocks, sched if K else None, K, native.ptr(self.wrench), slots,
            native.ptr(self.grid_ws) if slots else None, dt, flags, s), "grid_update")

    def _wrench_flush(self, s):
        if self.deterministic and self.n_slots:
            native.check(self.lib.mpm_grid_wrench_flush(
                self.pG, self.n_blocks, native.ptr(self.wrench), self.n_slots, native.ptr(self.grid_ws), s),
                "wrench_flush")

    def _enqueue_tick_cdf(self, dt, S, K, fused, lists, s):
        """One tick with the CPIC per-substep host work between the launches (the
        reference's p2g2p / p2g2p_fused_tick with _stamp_cdf and _integrate_cdf_poses,
        mpm_solver_warp.py:1130-1131, 1334-1336, 1410-1413, 1536-1539)."""
        G2P, ST, P2G = native.PASS_G2P, native.PASS_STRESS, native.PASS_P2G
        cdf = self.cdf
        times, _ = ColliderSet.times(self.time, dt, S)
        for k in range(S):
            t = float(times[k])
            cdf.stamp(t, lag_prev=fused)
            if self.modifiers and not fused:
                self._apply_modifiers(k, s)
            self._pass(G2P | ST | P2G if (fused and k > 0) else ST | P2G, dt, s)
            self._grid(dt, k, K, lists, s)
            cdf.integrate(t, dt, lag_prev=fused)
            if not fused:
                self._pass(G2P, dt, s)
                if self.n_rigid_bodies:
                    self._rigid(k, dt, s)
        if fused:
            cdf.reconcile_prev()
            self._pass(G2P, dt, s)
        self._wrench_flush(s)

    def _enqueue_tick(self, dt, S, K, fused, lists, s):
        G2P, ST, P2G = native.PASS_G2P, native.PASS_STRESS, native.PASS_P2G
        # the CPIC ghost reads the stored particle velocity: no lean passes with CDF
        lean = native.PASS_LEAN if (self.cdf is None or self.cdf.n == 0) else 0
        if fused:
            self._pass(ST | P2G, dt, s)
            self._grid(dt, 0, K, lists, s)
            for k in range(1, S):
                mode = G2P | ST | P2G | lean
                if k == S - 1 and lean:
                    mode |= native.PASS_KEEP_STRESS
                self._pass(mode, dt, s)
                self._grid(dt, k, K, lists, s)
            self._pass(G2P, dt, s)
            self._wrench_flush(s)
        else:
            for k in range(S):
                if self.modifiers:
                    self._apply_modifiers(k, s)
                self._pass(ST | P2G, dt, s)
                self._grid(dt, k, K, lists, s)
                self._pass(G2P, dt, s)
                if self.n_rigid_bodies:
                    self._rigid(k, dt, s)
            self._wrench_flush(s)

    def _enqueue_tick_profiled(self, dt, S, K, fused, lists):
        torch = self.torch
        s = self.stream()
        ev = []

        def timed(name, fn):
            a = torch.cuda.Event(enable_timing=True)
            b = torch.cuda.Event(enable_timing=True)
            a.record()
            fn()
            b.record()
            ev.append((name, a, b))

        G2P, ST, P2G = native.PASS_G2P, native.PASS_STRESS, native.PASS_P2G
        for k in range(S):
            if self.modifiers and not fused:
                timed("particle_modifiers", lambda k=k: self._apply_modifiers(k, s))
            if fused and k > 0:
                timed("g2p2g_fused", lambda: self._pass(G2P | ST | P2G, dt, s))
            else:
                timed("stress_p2g", lambda: self._pass(ST | P2G, dt, s))
            timed("grid_update+BC", lambda k=k: self._grid(dt, k, K, lists, s))
            if not fused:
                timed("g2p", lambda: self._pass(G2P, dt, s))
                if self.n_rigid_bodies:
                    timed("rigid_body", lambda k=k: self._rigid(k, dt, s))
        if fused:
            timed("g2p", lambda: self._pass(G2P, dt, s))
        timed("wrench_flush", lambda: self._wrench_flush(s))
        torch.cuda.synchronize(self.device)
        for name, a, b in ev:
            self.time_profile.setdefault(name, []).append(a.elapsed_time(b))

    # --------------------------------------------------------------- exports
    def export_R(self):
        torch = self.torch
        R = torch.empty(self.n_particles, 9, dtype=torch.float32, device=self.device)
        native.check(self.lib.mpm_export_R(self.pP, native.ptr(R), self.stream()), "export_R")
        return R

    def export_cov(self):
        torch = self.torch
        if self.update_cov_with_F:
            return self.cov.t()
        out = torch.empty(self.n_particles, 6, dtype=torch.float32, device=self.device)
        if self.init_cov is None:
            return out.zero_()
        native.check(self.lib.mpm_export_cov(self.pP, native.ptr(self.init_cov), native.ptr(out),
                                             self.stream()), "export_cov")
        return out

    def grid_dense(self):
        """Dense (n, n, n) mass and (n, n, n, 3) v_in / v_out copies (z fastest)."""
        torch = self.torch
        n = self.n_grid
        m = torch.empty(n, n, n, dtype=torch.float32, device=self.device)
        vi = torch.empty(n, n, n, 3, dtype=torch.float32, device=self.device)
        vo = torch.empty(n, n, n, 3, dtype=torch.float32, device=self.device)
        native.check(self.lib.mpm_grid_to_dense(self.pG, native.ptr(m), native.ptr(vi), native.ptr(vo),
                                                self.stream()), "grid_to_dense")
        return m, vi, vo

    def finalize_mu_lam_bulk(self):
        """mpm_solver_warp.py:808-809 (an alias of finalize_mu_lam)."""
        self.finalize_mu_lam()

    def reset_densities_and_update_masses(self, all_particle_densities, device="cuda:0"):
        """mpm_solver_warp.py:1622-1635: mass = density * vol per particle."""
        d = all_particle_densities.detach().to(device=self.device, dtype=self.mass.dtype).reshape(-1)
        if d.shape[0] != self.n_particles:
            raise ValueError(f"{d.shape[0]} densities for {self.n_particles} particles")
        self.density.copy_(d)  # particle_density first, as the reference (mpm_solver_warp.py:1624)
        self.torch.mul(self.density, self.vol, out=self.mass)
        self._m_max = None  # the fixed-point mass scale follows the heaviest particle

    # ------------------------------------------- torch imports / exports (the simulator's API)
    # MPM_Simulator_WARP's import_*_from_torch / export_*_to_torch (mpm_solver_warp.py:
    # 1638-1700, 1785-1860), same names, shapes and in-place copy semantics; `clone` and
    # `device` are accepted for signature compatibility (imports always copy).
    def _import_rows(self, dst, tensor, width, name):
        t = tensor.detach().to(device=self.device, dtype=dst.dtype).reshape(-1, width)
        if t.shape[0] != self.n_particles:
            raise ValueError(f"import_particle_{name}: {t.shape[0]} rows for {self.n_particles} particles")
        dst.copy_(t.t() if dst.dim() == 2 else t.reshape(-1))

    def import_particle_x_from_torch(self, tensor_x, clone=True, device="cuda:0"):
        if tensor_x is not None:
            self._import_rows(self.x, tensor_x, 3, "x")

    def import_particle_v_from_torch(self, tensor_v, clone=True, device="cuda:0"):
        if tensor_v is not None:
            self._import_rows(self.v, tensor_v, 3, "v")

    def import_particle_F_from_torch(self, tensor_F, clone=True, device="cuda:0"):
        if tensor_F is not None:  # (n, 3, 3) or (n, 9), row-major
            self._import_rows(self.F, tensor_F, 9, "F")

    def import_particle_C_from_torch(self, tensor_C, clone=True, device="cuda:0"):
        if tensor_C is not None:
            self._import_rows(self.C, tensor_C, 9, "C")

    def import_particle_selection_from_torch(self, tensor_selection, clone=True, device="cuda:0"):
        """Nonzero entries leave the particle out of every transfer (as the reference)."""
        if tensor_selection is not None:
            self._import_rows(self.selection, tensor_selection, 1, "selection")

    def import_particle_material_from_torch(self, tensor_material, clone=True, device="cuda:0"):
        if tensor_material is not None:
            self._import_rows(self.material, tensor_material, 1, "material")
            # the ids present select the particle-pass specialisation
            self._mat_ids = {int(m) for m in self.torch.unique(self.material).cpu().tolist()}

    def export_particle_material_to_torch(self):
        return self.material

    def export_particle_selection_to_torch(self):
        return self.selection

    def export_particle_x_to_torch(self):
        return self.x.t().contiguous()

    def export_particle_v_to_torch(self):
        return self.v.t().contiguous()

    def export_particle_vol_to_torch(self):
        return self.vol

    def export_particle_stress_to_torch(self):
        return self.stress.t().contiguous()

    def export_particle_F_to_torch(self):
        return self.F.t().contiguous()

    def export_particle_L_to_torch(self):
        return self.L.t().contiguous()

    def export_particle_C_to_torch(self):
        return self.C.t().contiguous()

    def export_particle_R_to_torch(self, device="cuda:0"):
        """compute_R_from_F's R^T per particle, (n, 9) row-major (as the reference stores it)."""
        return self.export_R()

    def export_particle_cov_to_torch(self, device="cuda:0"):
        """Flat (6 n,) upper triangles (xx, xy, xz, yy, yz, zz), as the reference's array."""
        return self.export_cov().reshape(-1)

    def export_grid_v_out_to_torch(self):
        return self.grid_dense()[2]

    def export_grid_v_in_to_torch(self):
        """The last scatter's momentum sums. The grid update consumes (clears) them, so after
        a completed substep this reads zero, where the reference keeps them until the next
        substep's zeroing; the deterministic scatter keeps its sums in the fixed-point
        buffer, so there it is zero throughout."""
        return self.grid_dense()[1]

    def export_grid_m_to_torch(self):
        """The last scatter's node masses, with the lifetime of export_grid_v_in_to_torch."""
        return self.grid_dense()[0]


@dataclass
class Solver:
    """warpmpm's Solver on the HIP backend. Device selection occurs at load time."""

    grid: GridConfig = field(default_factory=GridConfig)
    device: str = "auto"
    guard_interval: int = 1
    sparse: bool = False
    fused: bool = True
    sort_interval: int = 0
    profile: bool = False
    inversion_policy: str = "warn"
    periodic_x: bool = False
    # "block": the reference's sort key; "cell": the same order refined by stencil base
    # cell, so that the particle pass can sum same-cell particles before its LDS atomics
    sort_key: str = "block"
    # no counterpart in the reference: 64-bit fixed-point integer accumulation of the
    # scatter and the wrench, bit-reproducible from run to run (and faster than float
    # atomics on gfx950). None = on unless a CDF collider is added (its blocked-deposit
    # wrench needs the float path); True = on, CDF colliders refused; False = float atomics.
    deterministic: bool | None = None
    _sim: Any = field(default=None, init=False, repr=False)
    _step: int = field(default=0, init=False, repr=False)
    _tick: int = field(default=0, init=False, repr=False)

    def __post_init__(self) -> None:
        if self.inversion_policy not in {"warn", "raise", "nan"}:
            raise ValueError("inversion_policy must be 'warn', 'raise', or 'nan'")
        if self.sort_key not in {"block", "cell"}:
            raise ValueError("sort_key must be 'block' or 'cell'")

    # ------------------------------------------------------------------ setup
    def load_particles(self, pos: np.ndarray, vol: np.ndarray, cov: np.ndarray | None = None,
                       cov_mode: str = "step") -> "Solver":
        pos = np.asarray(pos)
        if pos.ndim != 2 or pos.shape[1] != 3 or len(pos) == 0:
            # (the reference fails on an empty set at its first guard, reducing an empty array)
            raise ValueError(f"load_particles: positions must be a non-empty (n, 3) array, got shape {pos.shape}")
        if np.asarray(vol).reshape(-1).shape[0] != len(pos):
            raise ValueError(f"load_particles: {np.asarray(vol).size} volumes for {len(pos)} particles")
        self.device = _resolve_device(self.device)
        cov6 = None
        if cov is not None:
            if cov_mode not in ("step", "from_F"):
                raise ValueError(f"cov_mode must be 'step' or 'from_F', got {cov_mode!r}")
            cov6 = _cov_to_six(cov)
            if cov6.shape[0] != len(pos):
                raise ValueError(f"cov has {cov6.shape[0]} rows, expected {len(pos)}")
        self._sim = HipSimulator(pos.astype(np.float32), vol, self.grid.n_grid, self.grid.grid_lim,
                                 self.device, cov=cov6)
        if self.periodic_x:
            self._sim.model[0]["periodic_x"] = 1
            self._sim._model_dirty = True
        if cov6 is not None and cov_mode == "step":
            self._sim.enable_cov_step()
        if self.deterministic is not False:
            self._sim.enable_deterministic()
        return self

    def set_material(self, material, **overrides: float) -> "Solver":
        if hasattr(material, "resolve"):
            name, params = material.resolve()
        else:
            name, params = str(material), {}
        self._sim.set_parameters_dict({"material": name, "g": [0.0, 0.0, -9.81], **params,
                                       **overrides})
        self._sim.finalize_mu_lam()
        return self

    def set_material_range(self, start: int, end: int, material, **overrides) -> "Solver":
        if hasattr(material, "resolve"):
            name, params = material.resolve()
        else:
            name, params = str(material), {}
        if name in ("tabulated_viscous", "tabulated_mu_i"):
            raise NotImplementedError(
                "tabulated materials are global-only (one eta table on the model); use set_material")
        self._sim.set_parameters_for_particles(start, end, {"material": name, **params, **overrides})
        return self

    def finalize_rigid_bodies(self) -> "Solver":
        """Compute each rigid body's centre of mass, mass and inertia from the particles
        assigned material "rigid" (grouped by obj_id in set_material_range); solver.py:
        193-198. Rigid scenes run the split pipeline; the body momentum sums are float
        atomics, so deterministic=None switches to the float scatter and
        deterministic=True refuses them."""
        if self.deterministic:
            raise NotImplementedError("the deterministic scatter does not cover rigid bodies")
        sim = self._sim
        if sim.deterministic:
            sim.disable_deterministic()
        sim.finalize_rigid_bodies()
        return self

    def rigid_state(self, body: int = 0) -> dict:
        """State of one rigid body: com (3,), v (3,), omega (3,), R (3, 3) (solver.py:200-210)."""
        return self._sim.rigid_state(body)

    # ------------------------------------------------------------------ colliders
    def _cs(self) -> ColliderSet:
        return self._sim.colliders

    def add_plane(self, point, normal, surface: str = "sticky", friction: float = 0.0,
                  restitution: float = 0.0) -> "Solver":
        """Plane collider (mpm_solver_warp.py:1880-2018). A restitution in (0, 1] also
        registers the plane as a rigid-body contact surface (:1915-1925)."""
        if restitution != 0.0 and not (0.0 < restitution <= 1.0):
            raise ValueError("restitution must be between 0 (exclusive) and 1 (inclusive).")
        if surface == "sticky" and friction != 0:
            raise ValueError("friction must be 0 on sticky surfaces.")
        point = [float(v) for v in point]
        scale = 1.0 / float(f32(math.sqrt(float(sum(float(x) ** 2 for x in normal)))))
        normal = [scale * float(x) for x in normal]
        if restitution != 0.0:
            self._sim.add_rigid_plane(point, normal, restitution, friction, 0.0, 999.0)
        c = Collider(native.COLLIDER_PLANE, "plane", 0.0, 999.0)
        c.point = [float(f32(v)) for v in point]
        c.normal = [float(f32(v)) for v in normal]
        c.surface = {"sticky": 0, "slip": 1, "cut": 11}.get(surface, 2)
        c.friction = float(f32(friction))
        axis = next((i for i in range(3) if abs(abs(normal[i]) - 1.0) < 1e-6), -1)
        if axis >= 0:
            c.aabb_kind = "plane"
            c.plane_axis = axis
            c.plane_offset = float(point[axis])
            c.plane_sign = float(normal[axis])
        else:
            c.label = "plane_free"
        self._cs().add(c)
        return self

    def add_box(self, center, half_size, velocity=(0.0, 0.0, 0.0), start_time: float = 0.0,
                end_time: float = 1.0e9) -> int:
        """Kinematic axis-aligned box (set_velocity_on_cuboid, mpm_solver_warp.py:2028-2106)."""
        c = Collider(native.COLLIDER_BOX, "cuboid_velocity", start_time, end_time)
        c.point = [float(f32(v)) for v in center]
        c.size = [float(f32(v)) for v in half_size]
        c.velocity = [float(f32(v)) for v in velocity]
        c.aabb_kind = "box"
        return self._cs().add(c, wrench=True)

    def set_box(self, handle: int, center=None, velocity=None) -> "Solver":
        c = self._cs().items[handle]
        if center is not None:
            c.point = [float(f32(v)) for v in center]
        if velocity is not None:
            c.velocity = [float(f32(v)) for v in velocity]
        return self

    def _wrench(self, handle: int):
        c = self._cs().items[handle]
        if c.slot < 0:
            raise ValueError(f"collider {handle} ({c.label}) has no reaction accumulator")
        return c

    def reset_tool_force(self, handle: int) -> "Solver":
        c = self._wrench(handle)
        self._sim.wrench[6 * c.slot:6 * c.slot + 6].zero_()
        return self

    def tool_force(self, handle: int, dt: float) -> np.ndarray:
        c = self._wrench(handle)
        return self._sim.wrench[6 * c.slot:6 * c.slot + 3].cpu().numpy().astype(float) / dt

    def add_cup(self, profile, center, quat=(1.0, 0.0, 0.0, 0.0), velocity=(0.0, 0.0, 0.0),
                omega=(0.0, 0.0, 0.0), friction: float = 0.05, sticky_cells: float = 1.5,
                contact_cells: float = 0.5, start_time: float = 0.0, end_time: float = 1.0e9) -> int:
        """Revolved-SDF glass (add_revolved_sdf_collider, mpm_solver_warp.py:2121-2269)."""
        from ..colliders.glass import quat_to_mat

        sticky_depth = min(sticky_cells * self.grid.dx, 0.45 * profile.wall_thickness)
        band = contact_cells * self.grid.dx
        c = Collider(native.COLLIDER_CUP, "revolved_cup", start_time, end_time)
        c.point = [float(f32(v)) for v in center]
        c.rot = np.asarray(quat_to_mat(quat), dtype=np.float64).reshape(3, 3).copy()
        c.velocity = [float(f32(v)) for v in velocity]
        c.omega = [float(f32(v)) for v in omega]
        c.friction = float(f32(friction))
        c.sticky_depth = float(f32(sticky_depth))
        c.contact_band = float(f32(band))
        c.shape = tuple(float(f32(v)) for v in (profile.outer_radius, profile.inner_radius,
                                                 profile.half_height, profile.inner_floor_z,
                                                 profile.fillet_radius))
        c.aabb_kind = "ball"
        c.ball_r = float(np.sqrt(profile.outer_radius ** 2 + profile.half_height ** 2)) + float(band)
        return self._cs().add(c, wrench=True)

    def set_cup(self, handle: int, center=None, quat=None, velocity=None, omega=None) -> "Solver":
        from ..colliders.glass import quat_to_mat

        c = self._cs().items[handle]
        if center is not None:
            c.point = [float(f32(v)) for v in center]
        if quat is not None:
            c.rot = np.asarray(quat_to_mat(quat), dtype=np.float64).reshape(3, 3).copy()
        if velocity is not None:
            c.velocity = [float(f32(v)) for v in velocity]
        if omega is not None:
            c.omega = [float(f32(v)) for v in omega]
        return self

    def reset_cup_wrench(self, handle: int) -> "Solver":
        c = self._wrench(handle)
        self._sim.wrench[6 * c.slot:6 * c.slot + 6].zero_()
        return self

    def cup_wrench(self, handle: int, dt: float) -> dict:
        c = self._wrench(handle)
        w = self._sim.wrench[6 * c.slot:6 * c.slot + 6].cpu().numpy().astype(float)
        return {"force": w[:3] / dt, "torque": w[3:] / dt}

    def add_domain_walls(self, start_time: float = 0.0, end_time: float = 1.0e9) -> "Solver":
        """Outward-velocity clamp within 3 cells of each face, as six face shells sharing
        one parameter block (add_bounding_box, mpm_solver_warp.py:2287-2380)."""
        n, pad = self.grid.n_grid, 3
        shells = [((0, 0, 0), (pad, n, n)), ((n - pad, 0, 0), (pad, n, n)),
                  ((0, 0, 0), (n, pad, n)), ((0, n - pad, 0), (n, pad, n)),
                  ((0, 0, 0), (n, n, pad)), ((0, 0, n - pad), (n, n, pad))]
        for box in shells:
            c = Collider(native.COLLIDER_WALLS, "domain_walls", start_time, end_time)
            c.aabb_kind = "fixed"
            c.fixed_box = box
            self._cs().add(c)
        return self

    def add_sdf_collider(self, sdf, center, quat=(0.0, 0.0, 0.0, 1.0), velocity=(0.0, 0.0, 0.0),
                         omega=(0.0, 0.0, 0.0), band=None, surface: str = "separable",
                         friction: float = 0.4, start_time: float = 0.0,
                         end_time: float = 1.0e9) -> int:
        """Mesh-SDF collider (mpm_solver_warp.py:2621-2777)."""
        torch = self._sim.torch
        vals = np.asarray(sdf.values)
        res = int(vals.shape[0])
        if band is None:
            band = float(self._sim.dx)
        sid = {"sticky": 0, "slip": 1, "separable": 2}[surface]
        bmin = float(min(vals[0].min(), vals[-1].min(), vals[:, 0, :].min(), vals[:, -1, :].min(),
                         vals[:, :, 0].min(), vals[:, :, -1].min()))
        if band >= bmin:
            raise ValueError(
                f"SDF collider contact band ({band * 1e3:.1f} mm) reaches or exceeds the SDF "
                f"grid margin ({bmin * 1e3:.1f} mm): near-surface space outside the stored grid "
                f"would get no collision. Rebuild the SDF with more margin_cells or reduce band.")
        ext = (res - 1) * float(sdf.cell)
        o = [float(v) for v in sdf.origin]
        corners = np.array([[o[0] + a * ext, o[1] + b * ext, o[2] + d * ext]
                            for a in (0, 1) for b in (0, 1) for d in (0, 1)])
        r_max = float(np.linalg.norm(corners, axis=1).max())
        c = Collider(native.COLLIDER_SDF, "mesh_sdf", start_time, end_time)
        val_d = torch.from_numpy(np.ascontiguousarray(vals, dtype=np.float32)).to(self._sim.device)
        grad_d = torch.from_numpy(np.ascontiguousarray(sdf.grads, dtype=np.float32)).to(self._sim.device)
        c.sdf = (val_d, grad_d, res, [float(f32(v)) for v in o], float(f32(sdf.cell)), r_max)
        c.point = [float(f32(v)) for v in center]
        c.quat = [float(f32(v)) for v in quat]
        c.velocity = [float(f32(v)) for v in velocity]
        c.omega = [float(f32(v)) for v in omega]
        c.band = float(f32(band))
        c.surface = sid
        c.friction = float(f32(friction))
        c.aabb_kind = "sdf"
        c.sdf_corners = corners
        c.guard = {"r_max": r_max, "band": float(band), "warned": False}
        return self._cs().add(c, wrench=True)

    def set_sdf_pose(self, handle: int, center=None, quat=None, velocity=None, omega=None) -> "Solver":
        c = self._cs().items[handle]
        g = c.guard
        if g is not None and not g["warned"]:
            jump = 0.0
            if center is not None:
                jump += float(np.linalg.norm(np.array([float(center[i]) - c.point[i] for i in range(3)])))
            if quat is not None:
                dot = abs(sum(float(quat[i]) * c.quat[i] for i in range(4)))
                jump += 2.0 * float(np.arccos(min(1.0, dot))) * g["r_max"]
            if jump > g["band"]:
                g["warned"] = True
                warnings.warn(
                    f"set_sdf_pose jumped the collider surface by up to {jump * 1e3:.2f} mm in "
                    f"one tick, more than the contact band ({g['band'] * 1e3:.2f} mm): material "
                    f"inside the swept region can be tunnelled through. Drive the pose in "
                    f"smaller per-tick steps (or via velocity/omega).", RuntimeWarning, stacklevel=2)
        if center is not None:
            c.point = [float(f32(v)) for v in center]
        if quat is not None:
            c.quat = [float(f32(v)) for v in quat]
        if velocity is not None:
            c.velocity = [float(f32(v)) for v in velocity]
        if omega is not None:
            c.omega = [float(f32(v)) for v in omega]
        return self

    def reset_sdf_force(self, handle: int) -> "Solver":
        return self.reset_cup_wrench(handle)

    def sdf_wrench(self, handle: int, dt: float) -> dict:
        return self.cup_wrench(handle, dt)

    def add_point_cloud_collider(self, points, padding: int = 1, start_time: float = 0.0,
                                 end_time: float = 1.0e9) -> "Solver":
        """Static occupancy collider (mpm_solver_warp.py:2545-2616)."""
        torch = self._sim.torch
        n = self.grid.n_grid
        idx = np.clip((np.asarray(points) * self._sim.inv_dx).astype(np.int32), 0, n - 1)
        occ = np.zeros((n, n, n), np.int32)
        occ[idx[:, 0], idx[:, 1], idx[:, 2]] = 1
        for _ in range(padding):
            d = occ.copy()
            d[1:] |= occ[:-1]; d[:-1] |= occ[1:]
            d[:, 1:] |= occ[:, :-1]; d[:, :-1] |= occ[:, 1:]
            d[:, :, 1:] |= occ[:, :, :-1]; d[:, :, :-1] |= occ[:, :, 1:]
            occ = d
        c = Collider(native.COLLIDER_POINTCLOUD, "pointcloud", start_time, end_time)
        c.occupancy = torch.from_numpy(occ).to(self._sim.device)
        a = np.argwhere(occ == 1)
        c.aabb_kind = "fixed"
        c.fixed_box = ((tuple(int(v) for v in a.min(0)),
                        tuple(int(a.max(0)[i] - a.min(0)[i] + 1) for i in range(3))) if len(a) else None)
        self._cs().add(c)
        return self

    def add_cdf_collider(self, cdf, center, quat=(0.0, 0.0, 0.0, 1.0), velocity=(0.0, 0.0, 0.0),
                         omega=(0.0, 0.0, 0.0), band=None, surface: str = "separable",
                         friction: float = 0.4, start_time: float = 0.0, end_time: float = 1.0e9) -> int:
        """CPIC thin-boundary collider from an open oriented surface (geometry.CDFData),
        solver.py:363-386: particle-node transfers are severed across the surface, so it
        holds material at zero wall thickness. Default band = min(built band, 2 dx).
        Handles are their own space (0..3)."""
        if self.periodic_x:
            raise NotImplementedError("CDF colliders do not wrap the periodic x axis")
        if self.deterministic:
            raise NotImplementedError("the deterministic scatter does not cover CDF colliders")
        sim = self._sim
        if sim.deterministic:  # deterministic=None: CDF scenes take the float-atomic scatter
            sim.disable_deterministic()
        return sim.cdf_colliders().add(
            cdf.values, cdf.valid, cdf.origin, cdf.cell, cdf.band, center, quat, velocity, omega,
            band, surface, friction, start_time, end_time, sim.time)

    def set_cdf_pose(self, handle: int, center=None, quat=None, velocity=None, omega=None) -> "Solver":
        """Start-of-tick pose of a CDF collider (solver.py:388-394); the pose advances
        with velocity / omega every substep."""
        self._sim.cdf.set_pose(handle, center=center, quat=quat, velocity=velocity, omega=omega)
        return self

    def reset_cdf_wrench(self) -> "Solver":
        """Zero every CDF lane's reaction accumulators (solver.py:396-399)."""
        if self._sim.cdf is not None:
            self._sim.cdf.reset_wrench()
        return self

    def cdf_wrench(self, handle: int, dt: float) -> dict:
        """Reaction the material exerts on a CDF collider since the last reset, from the
        blocked scatter deposits, divided by dt (solver.py:401-413)."""
        f, t = self._sim.cdf.wrench(handle)
        return {"force": f / dt, "torque": t / dt}

    # ------------------------------------------------------------------ stepping
    def step(self, dt: float, substeps: int = 1) -> "Solver":
        """solver.py:429-452: guard + launch box, optional block sort, optional sparse
        rebuild, then one control tick of `substeps` substeps."""
        sim = self._sim
        # collider rows first: the host builds them while the GPU finishes the last tick
        sim.prepare_tick(dt, substeps)
        try:
            if self._tick % max(1, self.guard_interval) == 0:
                self._update_grid_box(dt, substeps)
            if self.sort_interval and self._tick % self.sort_interval == 0:
                self._sort_particles()
            if self.sparse:
                sim.rebuild_active_blocks()
        except BaseException:
            sim.drop_prepared()
            raise
        sim.profile = self.profile
        self._tick += 1
        sim.run_tick(dt, substeps, fused=self.fused, sparse=self.sparse)
        self._step += substeps
        return self

    def _sort_particles(self) -> bool:
        return self._sim.sort_by_block(cells=self.sort_key == "cell")

    def _update_grid_box(self, dt: float, substeps: int) -> None:
        """solver.py:485-518 on device-reduced bounds."""
        lo_x, hi_x, vmax, bad, flag = self._sim.read_bounds()
        dx = self.grid.dx
        lim = self.grid.grid_lim
        if flag & 4:
            raise RuntimeError("internal: a solid material ran in the viscous-only particle pass")
        if flag & native.ERR_FIXED_RANGE:
            raise RuntimeError("deterministic scatter: a contribution exceeded the fixed-point range "
                               "(the state is far outside the expected momentum scale)")
        if flag & native.ERR_POOL:
            raise RuntimeError("a particle scattered into a block outside the active set during the last "
                               "tick (it moved more than a block within the tick)")
        if flag & 1:
            raise RuntimeError("a particle stencil left the grid during the last tick (out of "
                               "bounds P2G was skipped); enlarge grid_lim or add a wall collider")
        gmin = lo_x[1:].min() if self.periodic_x else lo_x.min()
        gmax = hi_x[1:].max() if self.periodic_x else hi_x.max()
        if bad or gmin < 1.5 * dx or gmax > lim - 2.5 * dx:
            raise RuntimeError(
                f"particles within 2 cells of the grid edge (x in [{gmin:.4f}, {gmax:.4f}] m, "
                f"domain [0, {lim}] m, dx={dx:.4f}): the P2G stencil would write out of bounds. "
                f"Enlarge grid_lim or add a bounding box / wall collider.")
        pad = 3.0 * dx + 1.5 * float(vmax) * dt * substeps
        lo = lo_x - pad
        hi = hi_x + pad
        if self.periodic_x:
            lo[0], hi[0] = 0.0, lim
        self._sim.grid_launch_box = self._sim._grid_box(lo, hi, halo=0)

    def profile_report(self) -> str:
        prof = self._sim.time_profile if self._sim is not None else {}
        rows = [(k, sum(v) / 1000.0, len(v)) for k, v in prof.items() if v]
        if not rows:
            return "profile_report: no samples (set solver.profile = True and step)"
        timed = sum(t for _, t, _ in rows)
        rows.sort(key=lambda r: -r[1])
        lines = [f"substep profile over {self._step} substeps (timed device work {timed:.3f}s; "
                 f"hipEvent-timed live launches):"]
        for name, tot, n in rows:
            lines.append(f"  {name:<28s} {tot:7.3f}s  {tot / max(n, 1) * 1000:8.3f} ms/launch"
                         f"  {tot / timed * 100:5.1f}%")
        return "\n".join(lines)

    # ------------------------------------------------------------------ imports / exports
    def set_x(self, pos: np.ndarray) -> "Solver":
        t = self._sim.torch.from_numpy(np.ascontiguousarray(pos, dtype=np.float32))
        self._sim.x.copy_(t.t())
        return self

    def set_v(self, vel: np.ndarray) -> "Solver":
        t = self._sim.torch.from_numpy(np.ascontiguousarray(vel, dtype=np.float32))
        self._sim.v.copy_(t.t())
        return self

    def x(self) -> np.ndarray:
        return self._sim.x.t().cpu().numpy()

    def v(self) -> np.ndarray:
        return self._sim.v.t().cpu().numpy()

    def F(self) -> np.ndarray:
        return self._sim.F.t().cpu().numpy().reshape(-1, 3, 3)

    def stress(self) -> np.ndarray:
        return self._sim.stress.t().cpu().numpy().reshape(-1, 3, 3)

    def L(self) -> np.ndarray:
        return self._sim.L.t().cpu().numpy().reshape(-1, 3, 3)

    def inverted_count(self) -> int:
        J = np.linalg.det(self.F())
        return int(np.count_nonzero(~np.isfinite(J) | (J <= 0.0)))

    def _checked_jacobian(self) -> np.ndarray:
        J = np.linalg.det(self.F())
        bad = ~np.isfinite(J) | (J <= 0.0)
        n = int(np.count_nonzero(bad))
        if not n:
            return J
        msg = (f"{n} particles have non-finite or non-positive det(F); current volume and "
               "Cauchy stress are undefined for an inverted configuration")
        if self.inversion_policy == "raise":
            raise RuntimeError(msg)
        if self.inversion_policy == "warn":
            if not getattr(self, "_warned_inverted", False):
                self._warned_inverted = True
                warnings.warn(f"{msg}; using |det(F)| for compatibility. Set inversion_policy="
                              "'raise' to reject or 'nan' to mark these particles.",
                              RuntimeWarning, stacklevel=3)
            out = np.abs(J)
            out[~np.isfinite(out)] = np.nan
            return out
        J = J.copy()
        J[bad] = np.nan
        return J

    @property
    def _vol0(self) -> np.ndarray:
        """Rest volumes in the current particle order (the reference's host copy, which it
        permutes at every sort, solver.py:466): read from the simulator's volume array,
        which the sort permutes on the device."""
        return self._sim.vol.cpu().numpy()

    def vol(self) -> np.ndarray:
        return self._vol0 * self._checked_jacobian()

    def box_contact_wrench(self, box_center, box_half, layer_cells: float | None = None,
                           compressive_only: bool = True) -> dict:
        """The reference's box_contact_wrench (coupling/wrench.py:15-46) evaluated on the
        device state: 8 floats come back instead of x, F and stress of every particle.
        Volume and Cauchy stress follow inversion_policy "warn" (|det F|, clipped at
        1e-9); with "raise" an inverted particle raises, with "nan" the host formula on
        the exported arrays is used (NaN volumes poison the layer thickness there)."""
        sim = self._sim
        if self.inversion_policy == "nan":
            from ..coupling.wrench import box_contact_wrench
            return box_contact_wrench(self.x(), self.cauchy(), self.vol(), box_center, box_half,
                                      layer_cells=layer_cells, dx=self.grid.dx,
                                      compressive_only=compressive_only)
        torch = sim.torch
        if getattr(sim, "_bcw_ws", None) is None:
            sim._bcw_ws = torch.zeros(int(sim.lib.mpm_box_contact_workspace_floats()), dtype=torch.float32,
                                      device=sim.device)
            sim._bcw_out = torch.zeros(8, dtype=torch.float32, device=sim.device)
        box = (C.c_float * 6)(*[float(v) for v in box_center], *[float(v) for v in box_half])
        layer = float(layer_cells) * self.grid.dx if layer_cells else 0.0
        native.check(sim.lib.mpm_box_contact_wrench(sim.pP, box, layer, int(bool(compressive_only)),
                                                    native.ptr(sim._bcw_out), native.ptr(sim._bcw_ws),
                                                    sim.stream()), "box_contact_wrench")
        out = sim._bcw_out.cpu().numpy().astype(np.float64)
        if out[7] > 0 and self.inversion_policy == "raise":
            raise RuntimeError(f"{int(out[7])} particles have non-finite or non-positive det(F); current "
                               "volume and Cauchy stress are undefined for an inverted configuration")
        n = int(out[6])
        return {"force": out[0:3].copy(), "torque": out[3:6].copy(), "n_contact": n if n >= 3 else 0,
                "Fz": float(out[2])}

    def cauchy(self) -> np.ndarray:
        J = self._checked_jacobian()
        if self.inversion_policy == "warn":
            J = np.clip(J, 1e-9, None)
        return self.stress() / J[:, None, None]

    def cov(self) -> np.ndarray:
        return self._sim.export_cov().cpu().numpy().reshape(-1, 6)

    def R(self) -> np.ndarray:
        R = self._sim.export_R().cpu().numpy().reshape(-1, 3, 3)
        return np.transpose(R, (0, 2, 1))

    def x_torch(self):
        return self._sim.x.t()

    def v_torch(self):
        return self._sim.v.t()

    def F_torch(self):
        return self._sim.F.t().reshape(-1, 3, 3)

    def cov_torch(self):
        return self._sim.export_cov().reshape(-1, 6)

    def R_torch(self):
        return self._sim.export_R().reshape(-1, 3, 3).transpose(1, 2)

    @property
    def n_particles(self) -> int:
        return 0 if self._sim is None else self._sim.n_particles
