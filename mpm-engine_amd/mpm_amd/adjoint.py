"""Differentiable rollouts: reverse-mode gradients of a Solver rollout on the GPU.

``Rollout(solver, dt, substeps)`` takes a configured single-scene ``Solver`` (particles
loaded, material set, colliders added) and exposes the rollout of `substeps` substeps as
a ``torch.autograd.Function`` of the per-particle Lame parameters, yield stress, initial
velocity and initial position, of the generalized-Newtonian fluid's parameters
(viscosity, bulk modulus, Herschel-Bulkley K and n of material 10) and of the local
mu(I) coefficients (mu_s, delta_mu, I0, grain diameter of material 9). It returns the
final positions and the accumulated collider wrench, so a loss on the final state and on
the plate reaction forces back-propagates to Young's modulus, Poisson's ratio, the yield
stress (through the per-particle parameters), the fluid's rheology, the granular
friction law and to the initial particle state. Materials: fixed-corotated jelly (0),
von Mises (1, with or without hardening: the hardened yield stress is carried as
particle state, its adjoint substep to substep), Drucker-Prager sand (2), local mu(I)
(9) and the generalized-Newtonian fluid (10).

Two ways to keep what the backward reads, chosen by the memory they need:

* tape (the default when it fits in half the free device memory): the forward runs the
  split passes (stress + P2G, grid update, G2P) and keeps every substep's particle state
  and grids; the backward is the hand-written adjoint kernels of libmpm_hip.so
  (mpm_adjoint_g2p, mpm_adjoint_grid, mpm_adjoint_p2g_stress) run backwards through the
  tape. Memory O(substeps x (particles + grid)), time one forward plus the adjoint kernels.
* checkpoints: the forward runs fused ticks of `checkpoint_every` substeps (the
  production substep), storing the particle state (x, v, C, F_trial, the velocity
  gradient L, which the viscous stress of the next substep reads, and the yield stress,
  which hardening changes) at every tick boundary. Backward, segment by segment from the
  last: restore the checkpoint, recompute the segment with the split passes while keeping
  every intermediate particle state and grid, then run the adjoint kernels backwards
  through the segment. Memory O(checkpoints + one segment), time about one forward for
  the recompute plus the adjoint kernels.

In both the grid kernels of a segment of `checkpoint_every` substeps cover a launch box
computed on the device from the particle bounds at the segment's start.

The reference solver has no gradient path; this module is the differentiation form
built on the same kernels and collider schedule as the forward Solver.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from .core import native

# Planes of a substep slot: x 3, the yield stress 1 (von Mises hardening changes it per
# substep), v 3, C 9, F_trial 9, F 9, stress 9, L 9. x and the yield stress lead, so the
# recompute carries them to the next slot with one copy (the passes write the rest).
_X, _YS, _V, _C, _FT, _F, _ST, _L0 = 0, 3, 4, 7, 16, 25, 34, 43
_SLOT_PLANES = 52
_CARRY = 4           # planes [0, 4): x and the yield stress
_CKPT_PLANES = 34    # a checkpoint holds a slot's planes [0, 25) (x, yield, v, C, F_trial) and L
_CK_L0 = 25
FLUID_ID = 10        # the material whose parameters form the `fluid` input
GRANULAR_ID = 9      # the material whose mu(I) coefficients form the `granular` input
_FLUID_FIELDS = ("plastic_viscosity", "bulk", "hardening", "softening")
_GRANULAR_FIELDS = ("muI_mu_s", "muI_delta_mu", "muI_I0", "muI_d")


def lame(E, nu):
    """compute_mu_lam_from_E_nu (mpm_utils.py:583-588) on torch tensors."""
    return E / (2.0 * (1.0 + nu)), E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu))


class Rollout:
    """Differentiable rollout of `substeps` substeps of dt on a configured Solver."""

    SUPPORTED_MATERIALS = frozenset((0, 1, 2, GRANULAR_ID, FLUID_ID))
    SUPPORTED_COLLIDERS = frozenset((native.COLLIDER_PLANE, native.COLLIDER_BOX, native.COLLIDER_WALLS,
                                     native.COLLIDER_POINTCLOUD, native.COLLIDER_CUP, native.COLLIDER_SDF))

    def __init__(self, solver, dt: float, substeps: int, checkpoint_every: int = 20, tape: bool | None = None):
        sim = solver._sim
        if sim is None:
            raise ValueError("load particles and set a material before building a Rollout")
        if sim.n_scenes != 1:
            raise ValueError("a Rollout differentiates one scene")
        if not sim._mat_ids <= self.SUPPORTED_MATERIALS:
            raise NotImplementedError(f"adjoint kernels cover material ids {sorted(self.SUPPORTED_MATERIALS)}; "
                                      f"the solver has {sorted(sim._mat_ids)}")
        kinds = {c.kind for c in sim.colliders.items}
        if not kinds <= self.SUPPORTED_COLLIDERS:
            raise NotImplementedError("adjoint kernels cover plane, box, wall, point-cloud, revolved-cup and "
                                      "mesh-SDF colliders")
        if sim.model[0]["periodic_x"]:
            raise NotImplementedError("periodic domains have no adjoint")
        if sim.cdf is not None and sim.cdf.n > 0:
            raise NotImplementedError("CPIC (CDF) colliders have no adjoint")
        if len(sim.colliders.items) > 16:
            raise ValueError("at most 16 colliders in a differentiable rollout")
        self.solver = solver
        self.sim = sim
        self.torch = sim.torch
        # Block-sort the particles once (Solver._sort_particles, solver.py:454-467) so the
        # adjoint scatter, like the forward pass, works on LDS tiles; inputs and outputs
        # keep the particle order the solver had when the Rollout was built.
        n0 = sim.n_particles
        self.perm = sim.torch.arange(n0, device=sim.device)
        if solver._sort_particles():
            self.perm = sim._last_perm.clone()
        sim.sync_chunk_capacity()  # the launches below keep this order for the whole rollout
        self.dt = float(dt)
        self.substeps = int(substeps)
        self.every = int(checkpoint_every)
        torch = self.torch
        dev = sim.device
        n = sim.n_particles
        self.n = n
        # initial state the rollout starts from (positions may be overridden per call)
        self.x0 = self._to_user(sim.x.t()).contiguous()
        self.C0 = sim.C.clone()
        self.Ft0 = sim.F_trial.clone()
        self.F0 = sim.F.clone()
        self.L0 = sim.L.clone()
        # the viscous stress reads L: slots carry it forward and checkpoints keep it
        self.viscous = FLUID_ID in sim._mat_ids
        md = sim.model[0]
        self.fluid0 = [float(md[fld][FLUID_ID]) for fld in _FLUID_FIELDS]
        self.granular0 = [float(md[fld][GRANULAR_ID]) for fld in _GRANULAR_FIELDS]
        # collider rows of every substep, integrated once from the current poses
        saved = [(list(c.point), np.array(c.rot), list(c.quat)) for c in sim.colliders.items]
        tab, _ = sim.colliders.schedule(sim.time, self.dt, self.substeps)
        for c, (p, r, q) in zip(sim.colliders.items, saved):
            c.point, c.rot, c.quat = p, r, q
        self.K = tab.shape[1]
        self.n_slots = sim.colliders.n_slots
        self.sched = torch.from_numpy(tab.view(np.uint8).reshape(-1).copy()).to(dev)
        self.row_bytes = self.K * native.COLLIDER_DTYPE.itemsize
        # launch box of the grid kernels, per checkpoint segment: the particle box at the
        # segment's start widened by 3 cells plus 1.5x the segment's travel at the largest
        # speed (the guard of Solver.step), computed on the device (mpm_bounds ->
        # mpm_box_from_bounds -> mpm_blocks_from_box) and kept for the backward
        nb = sim.n_blocks
        self.ctl = torch.zeros(native.TICK_CTL_DTYPE.itemsize // 4, dtype=torch.int32, device=dev)
        self.blocks = torch.zeros(nb, dtype=torch.int32, device=dev)  # the box's block list
        self.bnd = torch.zeros(9, dtype=torch.float32, device=dev)
        self.box_pad = 3.0 * float(sim.dx)
        self.ws = torch.zeros(max(int(sim.lib.mpm_grid_workspace_floats(nb, self.n_slots)), 1),
                              dtype=torch.float32, device=dev)
        self.wrench = torch.zeros(max(6 * self.n_slots, 6), dtype=torch.float32, device=dev)
        nseg = (self.substeps + self.every - 1) // self.every
        self.seg_ctl = torch.zeros(nseg, self.ctl.numel(), dtype=torch.int32, device=dev)
        node_floats = nb * 64 * 4
        if tape is None:  # the whole tape when it takes at most half the free device memory
            tape_bytes = 4 * ((self.substeps + 1) * _SLOT_PLANES * n + 2 * self.substeps * node_floats)
            tape = tape_bytes <= torch.cuda.mem_get_info(dev)[0] // 2
        self.tape = bool(tape)
        # particle slots and grids of every substep (tape) or of one segment (checkpoints)
        L = self.substeps if self.tape else min(self.every, self.substeps)
        self.ckpt = None if self.tape else torch.empty(nseg, _CKPT_PLANES, n, dtype=torch.float32, device=dev)
        self.slots = torch.zeros(L + 1, _SLOT_PLANES, n, dtype=torch.float32, device=dev)
        self.g_in = torch.zeros(L, node_floats, dtype=torch.float32, device=dev)
        self.g_out = torch.zeros(L, node_floats, dtype=torch.float32, device=dev)
        # adjoint buffers
        z = lambda *s: torch.zeros(*s, dtype=torch.float32, device=dev)  # noqa: E731
        self.adj = {"x": z(3, n), "v": z(3, n), "C": z(9, n), "F_trial": z(9, n), "F": z(9, n),
                    "stress": z(9, n), "mu": z(n), "lam": z(n), "yield_stress": z(n),
                    "grid_v": z(node_floats), "grid_mv": z(node_floats), "L": z(9, n),
                    "visc": z(n), "bulk": z(n), "hb_K": z(n), "hb_n": z(n),
                    "mui_mu_s": z(n), "mui_dmu": z(n), "mui_I0": z(n), "mui_d": z(n)}
        A = native.Adjoint()
        for k, t in self.adj.items():
            setattr(A, k, native.ptr(t))
        self.A = A
        self.impb = z(max(6 * self.n_slots, 6))
        # collider rows of the segment being reversed (a fixed buffer, so one captured
        # graph serves every full segment; the tape reads the whole schedule)
        self.seg_sched = None if self.tape else torch.zeros(max(L * self.row_bytes, 1), dtype=torch.uint8,
                                                             device=dev)
        # ABI structs of every substep slot, built once. The recompute's stress + P2G pass
        # of substep k reads slot k's state and L and writes F and stress into slot k + 1,
        # hardening slot k + 1's copy of the yield stress; its G2P pass moves slot k + 1's
        # copy of x and writes v, C, L and F_trial there.
        self._P_p2g = [self._particles(self.slots[k], self.slots[k + 1], self.slots[k], ys_slot=self.slots[k + 1])
                       for k in range(L)]
        self._P_g2p = [self._particles(self.slots[k + 1], self.slots[k + 1], ys_slot=self.slots[k + 1])
                       for k in range(L)]
        self._P_adj = [self._particles(self.slots[k], self.slots[k + 1]) for k in range(L)]
        # the stress adjoint of substep k reads L_k and the yield stress ys_k, which slot k holds
        self._P_st = [self._particles(self.slots[k], self.slots[k + 1], self.slots[k], ys_slot=self.slots[k])
                      for k in range(L)]
        self._G = [self._grid_struct(self.g_in[k], self.g_out[k]) for k in range(L)]
        self.use_graph = True
        self._graph = None     # checkpoints: one full segment's backward; tape: the whole backward
        self._fwd_graph = None  # tape: the forward's substep loop
        # checkpoints, slots and model fields hold the most recent forward only: each
        # forward gets a serial number and a backward refuses to run on another's data
        self.forward_count = 0

    # ------------------------------------------------------------------ structs
    def _grid_struct(self, node_in, node_out):
        G = native.Grid()
        C.memmove(C.byref(G), C.byref(self.sim.G), C.sizeof(G))
        G.node_in, G.node_out = native.ptr(node_in), native.ptr(node_out)
        G.block_touch = None  # the simulator's block state belongs to its own grid
        return G

    def _particles(self, state_slot, post_slot, l_slot=None, ys_slot=None):
        """Particle struct: x, v, C, F_trial from state_slot's planes, F and stress from
        post_slot's, L from l_slot's (default post_slot), the yield stress from ys_slot's
        (default the simulator's array); other per-particle arrays from the simulator."""
        P = native.Particles()
        C.memmove(C.byref(P), C.byref(self.sim.P), C.sizeof(P))
        n = self.n
        base = native.ptr(state_slot)
        post = native.ptr(post_slot)
        lp = native.ptr(post_slot if l_slot is None else l_slot)
        P.x, P.v, P.C, P.F_trial = base + _X * 4 * n, base + _V * 4 * n, base + _C * 4 * n, base + _FT * 4 * n
        P.F, P.stress, P.L = post + _F * 4 * n, post + _ST * 4 * n, lp + _L0 * 4 * n
        if ys_slot is not None:
            P.yield_stress = native.ptr(ys_slot) + _YS * 4 * n
        P.cov = 0
        return P

    # ------------------------------------------------------------------ ordering
    def _to_slots(self, t):
        """User particle order -> the rollout's internal (block-sorted) order."""
        return t[self.perm]

    def _to_user(self, t):
        out = self.torch.empty_like(t)
        out[self.perm] = t
        return out

    # ------------------------------------------------------------------ forward
    def _set_state(self, mu, lam, ys, v0, x0, fluid, granular):
        sim = self.sim
        md = sim.model[0]
        for fields, mat, vals in ((_FLUID_FIELDS, FLUID_ID, fluid), (_GRANULAR_FIELDS, GRANULAR_ID, granular)):
            for fld, val in zip(fields, vals.cpu().tolist()):
                md[fld][mat] = val
        sim._model_dirty = True
        sim.mu.copy_(self._to_slots(mu))
        sim.lam.copy_(self._to_slots(lam))
        sim.yield_stress.copy_(self._to_slots(ys))
        sim.x.copy_(self._to_slots(x0).t())
        sim.v.copy_(self._to_slots(v0).t())
        sim.C.copy_(self.C0)
        sim.F_trial.copy_(self.Ft0)
        sim.F.copy_(self.F0)
        sim.L.copy_(self.L0)
        sim.stress.zero_()
        sim._upload_model()

    def _grid_update(self, G, row, dt, stream, wrench, slots, flags):
        sim = self.sim
        sched = native.ptr(self.sched) + row * self.row_bytes
        native.check(sim.lib.mpm_grid_update(
            C.byref(G), native.ptr(sim.model_dev), native.ptr(self.ctl), native.ptr(self.blocks),
            native.ptr(self.ctl) + 24, sim.n_blocks, sched if self.K else None, self.K,
            native.ptr(wrench) if slots else None, slots, native.ptr(self.ws) if slots else None,
            dt, flags, stream), "grid_update")

    def forward(self, mu, lam, ys, v0, x0, fluid, granular):
        """Fused forward rollout from (x0, v0); returns (x_T (n, 3), wrench (slots, 6)).

        The adjoint reads the float scatter result (node_in), so the rollout runs the
        float-atomic scatter; a solver in the (default) fixed-point mode is switched back
        to it when the rollout returns."""
        sim = self.sim
        was_det = sim.deterministic
        if was_det:
            sim.disable_deterministic()
        try:
            if self.tape:
                return self._forward_tape(mu, lam, ys, v0, x0, fluid, granular)
            return self._forward(mu, lam, ys, v0, x0, fluid, granular)
        finally:
            self.forward_count += 1
            if was_det:
                sim.enable_deterministic()

    def _forward(self, mu, lam, ys, v0, x0, fluid, granular):
        sim = self.sim
        self._set_state(mu, lam, ys, v0, x0, fluid, granular)
        self.wrench.zero_()
        sim.error_flag.zero_()
        s = sim.stream()
        lib = sim.lib
        dt = self.dt
        G2P, ST, P2G = native.PASS_G2P, native.PASS_STRESS, native.PASS_P2G
        pG = sim.pG
        zl = (C.c_int32 * 3)(0, 0, 0)
        zh = (C.c_int32 * 3)(*([sim.n_grid] * 3))
        for seg, k0 in enumerate(range(0, self.substeps, self.every)):
            S = min(self.every, self.substeps - k0)
            ck = self.ckpt[seg]
            ck[_X:_X + 3].copy_(sim.x)
            ck[_YS].copy_(sim.yield_stress)  # hardened in place by the passes
            ck[_V:_V + 3].copy_(sim.v)
            ck[_C:_C + 9].copy_(sim.C)
            ck[_FT:_FT + 9].copy_(sim.F_trial)
            if self.viscous:
                ck[_CK_L0:_CK_L0 + 9].copy_(sim.L)
            native.check(lib.mpm_grid_zero_box(pG, zl, zh, 3, s), "zero_box")
            self._segment_box(seg, S, s)
            G = sim.G
            sim._pass(ST | P2G, dt, s)
            self._grid_update(G, k0, dt, s, self.wrench, self.n_slots, native.GRID_ZERO_IN)
            for k in range(1, S):
                mode = G2P | ST | P2G | native.PASS_LEAN
                if k == S - 1:
                    mode |= native.PASS_KEEP_STRESS
                sim._pass(mode, dt, s)
                self._grid_update(G, k0 + k, dt, s, self.wrench, self.n_slots, native.GRID_ZERO_IN)
            sim._pass(G2P, dt, s)
        flag = int(sim.error_flag.item())
        if flag & (native.ERR_STENCIL | native.ERR_NONFINITE):
            raise RuntimeError(f"forward rollout left the grid or produced non-finite state (flag {flag})")
        wrench = self.wrench[:6 * self.n_slots].clone().view(-1, 6)
        return self._to_user(sim.x.t()), wrench

    def _forward_tape(self, mu, lam, ys, v0, x0, fluid, granular):
        sim = self.sim
        torch = self.torch
        self._set_state(mu, lam, ys, v0, x0, fluid, granular)
        self.wrench.zero_()
        sim.error_flag.zero_()
        first = self.slots[0]
        first[_X:_X + 3].copy_(sim.x)
        first[_YS].copy_(sim.yield_stress)
        first[_V:_V + 3].copy_(sim.v)
        first[_C:_C + 9].copy_(sim.C)
        first[_FT:_FT + 9].copy_(sim.F_trial)
        first[_L0:_L0 + 9].copy_(sim.L)
        self.g_in.zero_()  # every P2G of the tape accumulates into its own grid
        if self.use_graph:
            if self._fwd_graph is None:
                g = torch.cuda.CUDAGraph()
                torch.cuda.synchronize(sim.device)
                with torch.cuda.graph(g):
                    self._tape_forward_body()
                self._fwd_graph = g
            self._fwd_graph.replay()
        else:
            self._tape_forward_body()
        # the simulator ends in the final state, as after the fused forward
        last = self.slots[self.substeps]
        sim.x.copy_(last[_X:_X + 3])
        sim.yield_stress.copy_(last[_YS])
        sim.v.copy_(last[_V:_V + 3])
        sim.C.copy_(last[_C:_C + 9])
        sim.F_trial.copy_(last[_FT:_FT + 9])
        sim.F.copy_(last[_F:_F + 9])
        sim.stress.copy_(last[_ST:_ST + 9])
        sim.L.copy_(last[_L0:_L0 + 9])
        flag = int(sim.error_flag.item())
        if flag & (native.ERR_STENCIL | native.ERR_NONFINITE):
            raise RuntimeError(f"forward rollout left the grid or produced non-finite state (flag {flag})")
        wrench = self.wrench[:6 * self.n_slots].clone().view(-1, 6)
        return self._to_user(sim.x.t()), wrench

    def _tape_forward_body(self):
        """Every substep with the split passes into its own slot and grids, segment by
        segment (each with its launch box from the positions at its start)."""
        sim = self.sim
        s = sim.stream()
        md = native.ptr(sim.model_dev)
        sched = native.ptr(self.sched)
        for seg, k0 in enumerate(range(0, self.substeps, self.every)):
            S = min(self.every, self.substeps - k0)
            self._segment_box(seg, S, s, C.byref(self._P_adj[k0]))
            for k in range(k0, k0 + S):
                self._substep(k, sched + k * self.row_bytes, s, md, True)

    def _substep(self, k, row, s, md, wrench):
        """Substep k with the split passes: slot k -> slot k + 1 through grids k (the
        collider rows at device address `row`; the wrench accumulated when `wrench`)."""
        sim, lib, dt = self.sim, self.sim.lib, self.dt
        self.slots[k + 1, :_CARRY].copy_(self.slots[k, :_CARRY])
        G = C.byref(self._G[k])
        native.check(lib.mpm_particle_pass(C.byref(self._P_p2g[k]), G, md, dt,
                                           native.PASS_STRESS | native.PASS_P2G, s), "stress + p2g")
        slots = self.n_slots if wrench else 0
        native.check(lib.mpm_grid_update(
            G, md, native.ptr(self.ctl), native.ptr(self.blocks), native.ptr(self.ctl) + 24,
            sim.n_blocks, row if self.K else None, self.K, native.ptr(self.wrench) if slots else None, slots,
            native.ptr(self.ws) if slots else None, dt, 0, s), "grid update")
        native.check(lib.mpm_particle_pass(C.byref(self._P_g2p[k]), G, md, dt, native.PASS_G2P, s), "g2p")

    def _reverse_substep(self, k, row, s, md):
        """The adjoint of substep k, from slots k, k + 1 and grids k."""
        lib, dt = self.sim.lib, self.dt
        P, G, pA = C.byref(self._P_adj[k]), C.byref(self._G[k]), C.byref(self.A)
        native.check(lib.mpm_adjoint_g2p(P, G, pA, md, dt, s), "adjoint g2p")
        native.check(lib.mpm_adjoint_grid(G, md, native.ptr(self.blocks), native.ptr(self.ctl) + 24,
                                          self.sim.n_blocks, row if self.K else None, self.K,
                                          native.ptr(self.impb), self.n_slots, pA, dt, s), "adjoint grid")
        native.check(lib.mpm_adjoint_p2g_stress(C.byref(self._P_st[k]), G, pA, md, dt, s), "adjoint p2g+stress")

    def _segment_box(self, seg, S, s, pP=None):
        """The segment's launch box from the particle positions (the simulator's, or those
        of the particle struct pP), on the device; kept for the segment's backward."""
        sim, lib = self.sim, self.sim.lib
        native.check(lib.mpm_bounds(sim.pP if pP is None else pP, native.ptr(self.bnd), s), "bounds")
        native.check(lib.mpm_box_from_bounds(native.ptr(self.bnd), sim.pG, self.box_pad, 1.5 * self.dt * S,
                                             native.ptr(self.ctl), s), "box_from_bounds")
        native.check(lib.mpm_blocks_from_box(sim.pG, native.ptr(self.ctl), native.ptr(self.blocks), s),
                     "blocks_from_box")
        self.seg_ctl[seg].copy_(self.ctl)

    # ------------------------------------------------------------------ backward
    def _load_segment(self, seg, k0, S):
        """Checkpoint of segment `seg` into slot 0, its collider rows into seg_sched and its
        launch box (and block list) into ctl."""
        self.slots[0, :_CK_L0].copy_(self.ckpt[seg, :_CK_L0])
        if self.viscous:
            self.slots[0, _L0:_L0 + 9].copy_(self.ckpt[seg, _CK_L0:_CK_L0 + 9])
        if self.K:
            nb = S * self.row_bytes
            self.seg_sched[:nb].copy_(self.sched[k0 * self.row_bytes:k0 * self.row_bytes + nb])
        self.ctl.copy_(self.seg_ctl[seg])
        sim = self.sim
        native.check(sim.lib.mpm_blocks_from_box(sim.pG, native.ptr(self.ctl), native.ptr(self.blocks),
                                                 sim.stream()), "blocks_from_box")

    def _segment_body(self, S):
        """Recompute S substeps from slot 0 with the split passes, keeping every state and
        grid, then run the adjoint kernels backwards through them."""
        sim = self.sim
        s = sim.stream()
        md = native.ptr(sim.model_dev)
        # P2G accumulates into g_in[k] (zero first); the grid update writes g_out[k] over
        # the segment's launch box, where every particle of the segment gathers
        self.g_in[:S].zero_()
        sched0 = native.ptr(self.seg_sched)
        for k in range(S):
            self._substep(k, sched0 + k * self.row_bytes, s, md, False)
        for k in reversed(range(S)):
            self._reverse_substep(k, sched0 + k * self.row_bytes, s, md)

    def _tape_backward_body(self):
        """The adjoint kernels backwards through the whole tape, segment by segment (each
        with the launch box its forward used)."""
        sim = self.sim
        s = sim.stream()
        md = native.ptr(sim.model_dev)
        sched = native.ptr(self.sched)
        for seg, k0 in reversed(list(enumerate(range(0, self.substeps, self.every)))):
            S = min(self.every, self.substeps - k0)
            self.ctl.copy_(self.seg_ctl[seg])
            native.check(sim.lib.mpm_blocks_from_box(sim.pG, native.ptr(self.ctl), native.ptr(self.blocks), s),
                         "blocks_from_box")
            for k in reversed(range(k0, k0 + S)):
                self._reverse_substep(k, sched + k * self.row_bytes, s, md)

    def backward(self, x_bar, wrench_bar, forward_id=None):
        """Adjoint of the last forward: returns (mu_bar, lam_bar, ys_bar, v0_bar, x0_bar,
        fluid_bar, granular_bar). forward_id (the forward's serial number, forward_count - 1
        when it returned) must name the last forward: the checkpoints belong to it."""
        if forward_id is not None and forward_id != self.forward_count - 1:
            raise RuntimeError("this Rollout ran another forward since the one being differentiated; "
                               "its checkpoints are gone (backward each result before the next forward, "
                               "or use one Rollout per differentiable call)")
        sim = self.sim
        was_det = sim.deterministic
        if was_det:
            sim.disable_deterministic()
        try:
            return self._backward(x_bar, wrench_bar)
        finally:
            if was_det:
                sim.enable_deterministic()

    def _backward(self, x_bar, wrench_bar):
        sim = self.sim
        torch = self.torch
        a = self.adj
        for t in a.values():
            t.zero_()
        a["x"].copy_(self._to_slots(x_bar).t())
        self.impb.zero_()
        if wrench_bar is not None and self.n_slots:
            self.impb[:6 * self.n_slots].copy_(wrench_bar.reshape(-1))
        sim._upload_model()
        if self.tape:
            if not self.use_graph:
                self._tape_backward_body()
            else:
                if self._graph is None:  # one capture of the whole backward
                    g = torch.cuda.CUDAGraph()
                    torch.cuda.synchronize(sim.device)
                    with torch.cuda.graph(g):
                        self._tape_backward_body()
                    self._graph = g
                self._graph.replay()
        else:
            for seg, k0 in reversed(list(enumerate(range(0, self.substeps, self.every)))):
                S = min(self.every, self.substeps - k0)
                self._load_segment(seg, k0, S)
                if self.use_graph and S == self.every:
                    if self._graph is None:  # one capture serves every full segment
                        g = torch.cuda.CUDAGraph()
                        torch.cuda.synchronize(sim.device)
                        with torch.cuda.graph(g):
                            self._segment_body(S)
                        self._graph = g
                    self._graph.replay()
                else:
                    self._segment_body(S)
        flag = int(sim.error_flag.item())
        if flag & (native.ERR_ADJ_MATERIAL | native.ERR_ADJ_COLLIDER):
            raise NotImplementedError(f"a material or collider without an adjoint was met (flag {flag})")
        fluid_bar = torch.stack([a[k].double().sum() for k in ("visc", "bulk", "hb_K", "hb_n")])
        granular_bar = torch.stack([a[k].double().sum() for k in ("mui_mu_s", "mui_dmu", "mui_I0", "mui_d")])
        return (self._to_user(a["mu"]), self._to_user(a["lam"]), self._to_user(a["yield_stress"]),
                self._to_user(a["v"].t()), self._to_user(a["x"].t()), fluid_bar, granular_bar)

    # ------------------------------------------------------------------ autograd
    def __call__(self, mu, lam, yield_stress, v0, x0=None, fluid=None, granular=None):
        """(x_T, wrench) as a differentiable function of per-particle mu, lam, yield
        stress (n,), initial velocity (n, 3), initial position (n, 3), the fluid
        parameters (4,) = (viscosity, bulk modulus, K, n) of material 10 and the mu(I)
        coefficients (4,) = (mu_s, delta_mu, I0, grain diameter) of material 9 (default:
        the solver's values)."""
        torch = self.torch
        if x0 is None:
            x0 = self.x0
        if fluid is None:
            fluid = torch.tensor(self.fluid0, dtype=torch.float32, device=self.sim.device)
        if granular is None:
            granular = torch.tensor(self.granular0, dtype=torch.float32, device=self.sim.device)
        return _RolloutFn.apply(self, mu, lam, yield_stress, v0, x0, fluid, granular)

    def _scalars(self, given, defaults):
        torch = self.torch
        dev = self.sim.device
        vals = [torch.as_tensor(defaults[i], dtype=torch.float32, device=dev) if v is None else v
                for i, v in enumerate(given)]
        return torch.stack([v.to(torch.float32).reshape(()) for v in vals])

    def run(self, E, nu, yield_stress, v0, x0=None):
        """Rollout as a function of the material's E, nu (scalars, broadcast to every
        particle through compute_mu_lam_from_E_nu) and yield stress (scalar or (n,))."""
        torch = self.torch
        mu, lam = lame(E, nu)
        ones = torch.ones(self.n, dtype=torch.float32, device=self.sim.device)
        ys = yield_stress * ones if yield_stress.dim() == 0 else yield_stress
        return self(mu * ones, lam * ones, ys, v0, x0)

    def run_fluid(self, yield_stress, v0, viscosity=None, bulk=None, K=None, n=None, x0=None):
        """Rollout as a function of the generalized-Newtonian fluid's viscosity, bulk
        modulus, Herschel-Bulkley K and n (scalars; None keeps the solver's value) and
        yield stress (scalar or (n,)); the Lame parameters keep the solver's values."""
        torch = self.torch
        fluid = self._scalars((viscosity, bulk, K, n), self.fluid0)
        ones = torch.ones(self.n, dtype=torch.float32, device=self.sim.device)
        ys = yield_stress * ones if yield_stress.dim() == 0 else yield_stress
        return self(self._to_user(self.sim.mu).clone(), self._to_user(self.sim.lam).clone(), ys, v0, x0, fluid)

    def run_granular(self, v0, mu_s=None, delta_mu=None, I0=None, grain_diameter=None, E=None, nu=None,
                     x0=None):
        """Rollout as a function of the local mu(I) coefficients mu_s, delta_mu, I0 and the
        grain diameter (scalars; None keeps the solver's value) and of E, nu (scalars
        broadcast through compute_mu_lam_from_E_nu; None keeps the solver's mu, lam)."""
        torch = self.torch
        granular = self._scalars((mu_s, delta_mu, I0, grain_diameter), self.granular0)
        ones = torch.ones(self.n, dtype=torch.float32, device=self.sim.device)
        if E is None and nu is None:
            mu, lam = self._to_user(self.sim.mu).clone(), self._to_user(self.sim.lam).clone()
        else:
            md = self.sim.model[0]
            E = torch.as_tensor(float(md["E"][GRANULAR_ID])) if E is None else E
            nu = torch.as_tensor(float(md["nu"][GRANULAR_ID])) if nu is None else nu
            m, l_ = lame(E, nu)
            mu, lam = m * ones, l_ * ones
        ys = self._to_user(self.sim.yield_stress).clone()
        return self(mu, lam, ys, v0, x0, None, granular)


def _as_f32(t, torch):
    return t.detach().to(torch.float32).contiguous()


def _make_fn():
    import torch

    class RolloutFn(torch.autograd.Function):
        @staticmethod
        def forward(ctx, ro, mu, lam, ys, v0, x0, fluid, granular):
            ctx.ro = ro
            ctx.dtypes = (mu.dtype, lam.dtype, ys.dtype, v0.dtype, x0.dtype, fluid.dtype, granular.dtype)
            xT, w = ro.forward(_as_f32(mu, torch), _as_f32(lam, torch), _as_f32(ys, torch),
                               _as_f32(v0, torch), _as_f32(x0, torch), _as_f32(fluid, torch),
                               _as_f32(granular, torch))
            ctx.forward_id = ro.forward_count - 1
            return xT, w

        @staticmethod
        def backward(ctx, xT_bar, w_bar):
            ro = ctx.ro
            if xT_bar is None:
                xT_bar = torch.zeros(ro.n, 3, dtype=torch.float32, device=ro.sim.device)
            mub, lamb, yb, vb, xb, fb, gb = ro.backward(_as_f32(xT_bar, torch),
                                                        None if w_bar is None else _as_f32(w_bar, torch),
                                                        ctx.forward_id)
            dts = ctx.dtypes
            return (None, mub.to(dts[0]), lamb.to(dts[1]), yb.to(dts[2]), vb.to(dts[3]), xb.to(dts[4]),
                    fb.to(dts[5]), gb.to(dts[6]))

    return RolloutFn


class _Lazy:
    fn = None

    @classmethod
    def apply(cls, *args):
        if cls.fn is None:
            cls.fn = _make_fn()
        return cls.fn.apply(*args)


_RolloutFn = _Lazy
