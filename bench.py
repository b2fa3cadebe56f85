"""Throughput of the explicit MLS-MPM/APIC substep on one MI355X.

Headline workload: the 192^3 Newtonian honey pour (340k particles at 8 per cell, two
revolved-SDF glasses, the source tilting at 1 rad/s about its rim, 432 substeps per
1/60 s frame). A "step" is one frame (one control tick of 432 substeps). Prints one JSON
line: value = particle-substeps/s over the timed frames (particles resident on the GPU),
the roofline of the dominant kernel (the fused g2p+stress+p2g particle pass, timed with
HIP events on its stream during the timed frames; HBM traffic from the rocprofv3 --pmc
summary under profiles/), and the CPU oracle's throughput on a bounded sample of the same
workload. Two more legs ride in the same line:
  "batched"  256 dough-press scenes (64^3 each, ~8.2M particles) in one launch per phase,
             100-substep hipGraph chunks, autotuned launch shape and sort interval, against
             the same scenes solved one after another with the plain Solver;
  "cdf_pour" the same pour with both glasses as CPIC thin-boundary sheets;
  "float_atomics" the pour with the float scatter (float atomics into the grid, as the
             reference accumulates; per-workgroup fixed-point sums in LDS) instead of the
             default fixed-point (bit-reproducible) one;
  "adjoint"  the squeeze identification (50k von Mises particles, 400 substeps): gradient
             of a loss on final positions and plate forces, backward time / forward time.

    python bench.py [--gpus 1] [--steps 10] [--warmup 2] [--no-cpu] [--legs pour,adjoint,batch,cdf,float]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "mpm-engine_amd"))
sys.path.insert(0, REPO)

HBM_PEAK_GBS = 8000.0
# algorithmic bytes (reference-state semantics): per particle-substep the fused pass reads
# x 12, F 36, selection 4, material 4, mass 4, vol 4, mu 4, lam 4, yield 4 (76 B) and the
# reference's kernel writes x 12, v 12, C 36, L 36, F_trial 36, F 36, stress 36 (204 B);
# per active grid node it writes the P2G result (16 B) and reads the G2P velocity (12 B).
# The grid update moves the remaining 56 B per active node (zero m+v_in 16, normalise
# read 16 + write 12, zero v_out 12): 280 B/particle + 84 B/active node per substep.
PASS_B_PER_PARTICLE = 280.0
PASS_B_PER_NODE = 28.0
GRID_B_PER_NODE = 56.0
# The three byte counts of the fused pass, per particle and per active node:
#   A (compulsory): only x and F carry state between fused substeps: the 76 B of reads
#     above plus the x and F writes (48 B);
#   B (reference-state semantics, the primary count): as above;
#   C (stencil touch, uncached): tier B's particle bytes plus, per particle, 27 node reads
#     of the gather (12 B) and 27 read-modify-writes of the scatter (m + mv, 2 x 16 B),
#     the traffic of a pass with no tile reuse; the grid is inside those touches.
TIERS = {"A": (124.0, 28.0), "B": (PASS_B_PER_PARTICLE, PASS_B_PER_NODE), "C": (280.0 + 27 * (12 + 32), 0.0)}
# measured ceiling of 64-bit global atomic adds on MI355X (a tile-flush pattern: each
# workgroup adds 4 channels into 256 consecutive nodes), for the atomic-throughput figure
ATOMIC_CEILING_GPS = 45.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-substeps", type=int, default=6)
    ap.add_argument("--no-graph", action="store_true", help="eager launches instead of hipGraph replays")
    ap.add_argument("--sort-interval", type=int, default=1)
    ap.add_argument("--sort-key", choices=("block", "cell"), default="cell")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-scenes leg")
    ap.add_argument("--batch-scenes", type=int, default=256)
    ap.add_argument("--batch-ticks", type=int, default=3, help="timed 100-substep graph chunks")
    ap.add_argument("--seq-scenes", type=int, default=0,
                    help="scenes timed one by one for the sequential baseline (0 = all)")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint leg")
    ap.add_argument("--legs", default="pour,adjoint,batch,cdf,float",
                    help="comma list of legs to run (pour, adjoint, batch, cdf, float); for profiling one leg")
    ap.add_argument("--float-frames", type=int, default=2, help="timed frames of the float-atomic pour")
    ap.add_argument("--cdf-frames", type=int, default=2, help="timed frames of the CDF-glass pour")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a device copy of 256 MiB (the byte-counter calibration for --pmc)")
    ap.add_argument("--no-traffic", action="store_true",
                    help="skip the rocprofv3 counter runs that measure the fused pass's HBM bytes and atomics")
    ap.add_argument("--traffic-child", action="store_true",
                    help="internal: the copy calibration, then the pour up to its held frame (run under rocprofv3)")
    return ap.parse_args()


def cpu_baseline(substeps: int):
    """The oracle (single-threaded C restatement of the reference kernels, driven by the
    restated host schedule) on the same pour, for `substeps` substeps of one frame."""
    from mpm_amd.scenes import PourScene
    from oracle.sim import OracleSolver

    sc = PourScene()
    s = sc.build(OracleSolver(sc.grid))
    sc.substeps_saved = sc.substeps
    t0 = time.perf_counter()
    s.set_cup(sc.src, center=sc.center, quat=sc.quat, velocity=[0, 0, 0], omega=sc.omega)
    s.step(sc.dt, substeps)
    el = time.perf_counter() - t0
    n = len(sc.pos)
    out = {"value": n * substeps / el, "unit": "particle-substeps/s", "cores": 1, "kind": "port",
           "sample": f"honey pour 192^3, {n} particles, {substeps} substeps of one frame "
                     f"({el:.1f} s, single-threaded C oracle)"}
    try:  # the same oracle built with OpenMP, on every host core this process may use
        out["threaded"] = cpu_baseline_threaded(4 * substeps)
    except Exception as e:  # reported beside the single-threaded figure
        out["threaded"] = {"error": f"{type(e).__name__}: {e}"}
    for name, fn in (("press_C3", cpu_press), ("batch_sequential_C4", cpu_batch_sequential),
                     ("squeeze_forward_C5", cpu_squeeze_forward)):
        try:
            out[name] = fn()
        except Exception as e:  # reported beside the headline figure
            out[name] = {"error": f"{type(e).__name__}: {e}"}
    return out


def cpu_press(substeps: int = 2):
    """The oracle (one thread) on the dough + sand press: 128^3, sparse blocks, ~597k
    particles of Herschel-Bulkley dough and mu(I) sand, a pressing box."""
    from mpm_amd.scenes import DoughSandPress
    from oracle.sim import OracleSolver

    sc = DoughSandPress()
    s = sc.build(OracleSolver(sc.grid, sparse=True, sort_interval=20))
    t0 = time.perf_counter()
    s.step(sc.dt, substeps)
    el = time.perf_counter() - t0
    n = len(sc.pos)
    return {"value": n * substeps / el, "unit": "particle-substeps/s", "cores": 1, "kind": "port",
            "sample": f"dough + sand press 128^3, {n} particles, {substeps} substeps ({el:.1f} s)"}


def cpu_batch_sequential(scenes=(0, 64, 128, 192), substeps: int = 5):
    """The oracle (one thread) solving dough-press scenes of the 256-scene batch one after
    another, as scene-substeps/s (the batched leg's sequential baseline on the host)."""
    from mpm_amd.scenes import DoughPressBatch, herschel_bulkley_dough
    from oracle.sim import OracleSolver

    sc = DoughPressBatch(n_scenes=256, seed=0)
    el, parts = 0.0, 0
    for k in scenes:
        s = OracleSolver(sc.grid)
        s.load_particles(sc.pos[k], sc.vol[k])
        s.set_material(herschel_bulkley_dough())
        s._sim.yield_stress[:] = sc.yield_stress[k]
        s.add_plane((0, 0, sc.floor), (0, 0, 1), "sticky")
        s.add_domain_walls()
        s.add_box(sc.box_center[k], sc.box_half[k], (0.0, 0.0, -sc.box_speed[k]))
        t0 = time.perf_counter()
        s.step(sc.dt, substeps)
        el += time.perf_counter() - t0
        parts += len(sc.pos[k])
    return {"value": len(scenes) * substeps / el, "unit": "scene-substeps/s", "cores": 1, "kind": "port",
            "sample": f"scenes {list(scenes)} of the 256-scene dough-press batch ({parts} particles), "
                      f"{substeps} substeps each ({el:.1f} s)"}


def cpu_squeeze_forward(substeps: int = 10):
    """The oracle (one thread) on the squeeze workload's forward rollout (50,653 von
    Mises particles between closing plates); a central-difference gradient costs two
    such rollouts per parameter direction."""
    from mpm_amd.scenes import SqueezeScene
    from oracle.sim import OracleSolver

    sc = SqueezeScene()
    s = sc.build(OracleSolver(sc.grid))
    t0 = time.perf_counter()
    s.step(sc.dt, substeps)
    el = time.perf_counter() - t0
    n = len(sc.pos)
    return {"value": n * substeps / el, "unit": "particle-substeps/s", "cores": 1, "kind": "port",
            "forward_s_per_400_substeps": round(el / substeps * 400, 1),
            "sample": f"squeeze 64^3, {n} particles, {substeps} substeps ({el:.1f} s)"}


def cpu_baseline_threaded(substeps: int):
    """The oracle's OpenMP build (particle loops over all threads, atomic scatter) on the
    same pour; threads = OMP_NUM_THREADS (or the CPU count)."""
    from mpm_amd.scenes import PourScene
    from oracle import sim as osim

    threads = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)
    model = ""
    try:
        with open("/proc/cpuinfo") as fh:
            model = next((ln.split(":", 1)[1].strip() for ln in fh if ln.startswith("model name")), "")
    except OSError:
        pass
    osim.use_threaded_build(True)
    try:
        sc = PourScene()
        s = sc.build(osim.OracleSolver(sc.grid))
        t0 = time.perf_counter()
        s.set_cup(sc.src, center=sc.center, quat=sc.quat, velocity=[0, 0, 0], omega=sc.omega)
        s.step(sc.dt, substeps)
        el = time.perf_counter() - t0
    finally:
        osim.use_threaded_build(False)
    n = len(sc.pos)
    return {"value": n * substeps / el, "unit": "particle-substeps/s", "cores": threads,
            "cpu_model": model, "kind": "port",
            "sample": f"honey pour 192^3, {n} particles, {substeps} substeps of one frame ({el:.1f} s, "
                      f"OpenMP build of the C oracle)"}


def batch_leg(args):
    """Batched dough-press scenes (one launch per phase over all scenes, 100-substep
    hipGraph chunks, autotuned launch shape and sort interval) against the same scenes
    solved one after another with the plain Solver on the same GPU."""
    import numpy as np
    import torch

    from mpm_amd import Solver
    from mpm_amd.batch import BatchSolver
    from mpm_amd.scenes import DoughPressBatch

    chunk = 100
    sc = DoughPressBatch(n_scenes=args.batch_scenes, seed=0)
    b = sc.build_batch(BatchSolver(grid=sc.grid, device="cuda:0"))
    ns, n = b.n_scenes, b.n_particles
    print(f"[batch] {ns} scenes, {n} particles; warm-up + autotune", file=sys.stderr, flush=True)
    b.step(sc.dt, chunk)
    tune = b.autotune(sc.dt, chunk, ticks=1)
    # hipGraph capture with the chosen shape, then one more chunk so that the launch
    # capacities (sized from the previous sort's chunk counts) have settled
    b.step(sc.dt, chunk)
    b.step(sc.dt, chunk)
    # at least one whole sort cycle is timed
    ticks = max(args.batch_ticks, b.sort_interval)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(ticks):
        b.step(sc.dt, chunk)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    if not np.isfinite(b.x(0)).all():
        raise RuntimeError("non-finite batch state")
    batched = ns * chunk * ticks / el
    sim = b._sim
    # grid storage: the shared block pool against dense per-scene node arrays (node_in and
    # node_out float4, node_fix 4 x int64 per node, for every block of every scene)
    node_bytes = 16 + 16 + (32 if sim.node_fix is not None else 0)
    grid = {"grid_bytes": sim.grid_bytes(), "pool_blocks": sim.pool_blocks,
            "active_blocks": int(sim._blk["counts"][0].item()) if sim._blk is not None else None,
            "blocks_dense": sim.n_blocks,
            "grid_bytes_if_dense": sim.grid_bytes() + sim.n_blocks * 64 * node_bytes
            - (sim.node_in.numel() * 4 + sim.node_out.numel() * 4
               + (sim.node_fix.numel() * 8 if sim.node_fix is not None else 0))}
    # sequential baseline: each scene alone in a plain Solver, one timed graph-replayed
    # chunk after two warm-up chunks, under two sort policies (the reference's default,
    # no sort, and the batch's autotuned one); the speedup is against the faster
    picks = range(ns) if not args.seq_scenes else np.linspace(0, ns - 1, args.seq_scenes).astype(int)
    print(f"[batch] batched {batched:.1f} scene-steps/s; sequential baseline over {len(picks)} scenes",
          file=sys.stderr, flush=True)
    seq = {}
    for policy, kw in (("no_sort", {}), ("batch_sort", {"sort_interval": b.sort_interval, "sort_key": "cell"})):
        seq_el, seq_steps = 0.0, 0
        for k, s_id in enumerate(picks):
            s = sc.build_one(Solver(grid=sc.grid, device="cuda:0", **kw), int(s_id))
            s.step(sc.dt, chunk)
            s.step(sc.dt, chunk)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            s.step(sc.dt, chunk)
            torch.cuda.synchronize()
            seq_el += time.perf_counter() - t0
            seq_steps += chunk
            del s
            if k % 64 == 63:
                print(f"[batch] sequential {policy} {k + 1}/{len(picks)}", file=sys.stderr, flush=True)
        seq[policy] = seq_steps / seq_el
    seq_rate = max(seq.values())
    return {"metric": "batched scene-steps/s", "value": round(batched, 1), "unit": "scene-substeps/s",
            "sequential_value": round(seq_rate, 1), "speedup_vs_sequential": round(batched / seq_rate, 2),
            "sequential_by_sort_policy": {k: round(v, 1) for k, v in seq.items()},
            "particle_substeps_per_s": round(n * chunk * ticks / el, 1),
            "ms_per_chunk": round(el / ticks * 1e3, 2), "timed_chunks": ticks,
            "config": {"workload": f"{ns} dough-press scenes, 64^3 each, Herschel-Bulkley K=50 n=0.5, "
                                   f"yield log-U[50,500] Pa per particle, box plate U[0.02,0.1] m/s",
                       "particles": n, "substeps_per_chunk": chunk, "dt": sc.dt,
                       "sequential_scenes_timed": len(picks)},
            "grid_storage": grid,
            "autotune": tune["chosen"], "autotune_shapes_s": tune["shapes"]}


def adjoint_leg(args):
    """Squeeze identification (64^3, 37^3 von Mises particles, two closing plates, 400
    substeps of 1e-4, launch boxes per 20 substeps): gradient of |x_T - x*|^2 + |plate
    impulses - target|^2 (target: yield x 1.1) with respect to E, nu, yield and v0, and the
    times of the differentiable forward and of the backward against the plain forward
    rollout on the same GPU; the whole substep tape kept (the default at this size), and
    the checkpointed rollout beside it."""
    import numpy as np
    import torch

    from mpm_amd import Solver
    from mpm_amd.adjoint import Rollout
    from mpm_amd.scenes import SqueezeScene

    sc = SqueezeScene()
    S, every = 400, 20
    dev = torch.device("cuda:0")
    # plain forward: the production Solver, 20 ticks of 20 substeps (graph replays)
    plain = sc.build(Solver(grid=sc.grid, device="cuda:0"))
    plain.step(sc.dt, every)
    plain.step(sc.dt, every)
    plain = sc.build(Solver(grid=sc.grid, device="cuda:0"))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(S // every):
        plain.step(sc.dt, every)
    torch.cuda.synchronize()
    t_plain = time.perf_counter() - t0
    del plain
    f64 = lambda v, g=False: torch.tensor(v, dtype=torch.float64, device=dev, requires_grad=g)  # noqa: E731
    E0, nu0, y0 = 2.0e5, 0.3, 1.0e3

    def timed(tape):
        s = sc.build(Solver(grid=sc.grid, device="cuda:0"))
        n = s.n_particles
        ro = Rollout(s, sc.dt, S, checkpoint_every=every, tape=tape)
        with torch.no_grad():
            xt, wt = ro.run(f64(E0), f64(nu0), f64(1.1 * y0), torch.zeros(n, 3, device=dev))
        x_star, w_star = xt.double(), wt[:, :3].double()
        times = []
        for rep in range(2):  # the first pass warms allocations and graph captures up
            E, nu, ys = f64(E0, True), f64(nu0, True), f64(y0, True)
            v0 = torch.zeros(n, 3, device=dev, requires_grad=True)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            xT, w = ro.run(E, nu, ys, v0)
            loss = ((xT.double() - x_star) ** 2).sum() + ((w[:, :3].double() - w_star) ** 2).sum()
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            loss.backward()
            torch.cuda.synchronize()
            t2 = time.perf_counter()
            times.append((t1 - t0, t2 - t1))
        if not all(np.isfinite(t.grad.item()) for t in (E, nu, ys)):
            raise RuntimeError("non-finite adjoint gradient")
        return n, times[-1], loss, (E, nu, ys, v0)

    # the whole tape (this rollout needs ~7.6 GB of the 288 GB) and, for comparison, the
    # checkpointed rollout (fused forward ticks, segments recomputed in the backward)
    n, (t_fwd, t_bwd), loss, (E, nu, ys, v0) = timed(True)
    _, (c_fwd, c_bwd), _, _ = timed(False)
    return {"metric": "adjoint/forward time", "value": round(t_bwd / t_plain, 3), "unit": "ratio",
            "higher_is_better": False,
            "mode": "tape", "backward_s": round(t_bwd, 4), "taped_forward_s": round(t_fwd, 4),
            "plain_forward_s": round(t_plain, 4),
            "forward_plus_backward_over_plain": round((t_fwd + t_bwd) / t_plain, 3),
            "checkpoints": {"backward_s": round(c_bwd, 4), "checkpointed_forward_s": round(c_fwd, 4),
                            "backward_over_plain": round(c_bwd / t_plain, 3),
                            "forward_plus_backward_over_plain": round((c_fwd + c_bwd) / t_plain, 3)},
            "loss": float(loss), "grad": {"E": E.grad.item(), "nu": nu.grad.item(), "yield": ys.grad.item(),
                                          "v0_norm": float(v0.grad.norm())},
            "config": {"workload": "squeeze identification: 64^3, von Mises E=2e5 nu=0.3 yield=1e3, "
                                   "two box plates closing at 0.05 m/s, target rollout at yield x 1.1",
                       "particles": n, "substeps": S, "dt": sc.dt, "checkpoint_every": every}}


COUNTER_SETS = (("FETCH_SIZE", "TCC_EA0_ATOMIC_sum"), ("WRITE_SIZE", "TCC_ATOMIC_sum"))
FUSED_KERNEL = "k_particle_pass<15,"  # G2P | STRESS | P2G | LEAN, the interior substep


def _counter_run(args, counters, outdir):
    """One rocprofv3 --pmc run of this script in --traffic-child mode; returns (per-dispatch
    rows as {name, grid, counters}, child's JSON line)."""
    import csv
    import subprocess

    cmd = ["rocprofv3", "--pmc", *counters, "-f", "csv", "-d", outdir, "-o", "t", "--",
           sys.executable, os.path.abspath(__file__), "--traffic-child", "--steps", str(args.steps),
           "--warmup", str(args.warmup), "--sort-interval", str(args.sort_interval), "--sort-key", args.sort_key]
    if args.no_graph:
        cmd.append("--no-graph")
    env = dict(os.environ, TMPDIR="/tmp")
    r = subprocess.run(["timeout", "-s", "KILL", "600", *cmd], cwd="/tmp", env=env, capture_output=True, text=True)
    line = next((ln for ln in reversed(r.stdout.splitlines()) if ln.startswith("{")), None)
    if r.returncode != 0 or line is None:
        raise RuntimeError(f"rocprofv3 run failed (rc {r.returncode}): {r.stderr.strip()[-300:]}")
    rows = {}
    with open(os.path.join(outdir, "t_counter_collection.csv")) as fh:
        for row in csv.DictReader(fh):
            d = rows.setdefault(int(row["Dispatch_Id"]), {"name": row["Kernel_Name"], "grid": int(row["Grid_Size"]),
                                                           "c": {}})
            d["c"][row["Counter_Name"]] = float(row["Counter_Value"])
    return [rows[k] for k in sorted(rows)], json.loads(line)


def measured_traffic(args, launches: int, avg_s: float, digest: str):
    """HBM bytes and global atomics per fused-pass launch, from two rocprofv3 --pmc runs
    (FETCH_SIZE + TCC_EA0_ATOMIC_sum, then WRITE_SIZE + TCC_ATOMIC_sum) of a child that
    replays this workload to the same state (the fixed-point scatter is bit-reproducible:
    the child's final-position digest must equal this run's) and counts the fused passes
    of the same held frame. Byte counters are scaled by what they read for a device copy
    of 256 MiB run first in the same child."""
    import shutil
    import tempfile

    if shutil.which("rocprofv3") is None:
        return None, "rocprofv3 not found"
    out = {}
    child = None
    for counters in COUNTER_SETS:
        d = tempfile.mkdtemp(prefix="mpm_pmc_", dir="/tmp")
        try:
            rows, child = _counter_run(args, counters, d)
        finally:
            shutil.rmtree(d, ignore_errors=True)
        copies = [r for r in rows if "elementwise" in r["name"] and r["grid"] >= (1 << 22)][:3]
        # the held frame's interior passes (its last one keeps the stress: another mode)
        fused = [r for r in rows if FUSED_KERNEL in r["name"].replace(" ", "")][-(launches - 1):]
        if len(copies) < 3 or not fused:
            return None, f"counter rows not found ({len(copies)} copies, {len(fused)} fused)"
        for cn in counters:
            out[cn] = float(np.mean([r["c"][cn] for r in fused]))
            if cn in ("FETCH_SIZE", "WRITE_SIZE"):
                out[cn + "_copy"] = float(np.mean([r["c"][cn] for r in copies]))
    copy_kib = 256 * 1024.0  # bytes read (and written) by one copy, in KiB
    f_scale, w_scale = copy_kib / out["FETCH_SIZE_copy"], copy_kib / out["WRITE_SIZE_copy"]
    read_b = out["FETCH_SIZE"] * 1024.0 * f_scale
    write_b = out["WRITE_SIZE"] * 1024.0 * w_scale
    atom = out["TCC_EA0_ATOMIC_sum"]
    return {"bytes": int(read_b + write_b), "read_bytes": int(read_b), "write_bytes": int(write_b),
            "fetch_scale": round(f_scale, 3), "write_scale": round(w_scale, 3),
            "atomics_per_launch": int(atom), "atomics_at_l2_per_launch": int(out["TCC_ATOMIC_sum"]),
            "atomic_G_per_s": round(atom / avg_s / 1e9, 2),
            "atomic_ceiling_G_per_s": ATOMIC_CEILING_GPS,
            "same_state": child.get("digest") == digest, "launches_counted": launches}, None


def copy_calibration():
    import torch

    n = 64 * 1024 * 1024  # 256 MiB of float32
    src = torch.ones(n, device="cuda:0")
    dst = torch.empty_like(src)
    for _ in range(3):
        torch.mul(src, 1.0, out=dst)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes_read": 4 * n, "bytes_written": 4 * n, "launches": 3}))


class FenceFreeEvent:
    """A HIP timing event created with hipEventDisableSystemFence: recording it does not
    write back and invalidate the caches (the default event's system-scope fence does,
    which slows the kernel after it), so event pairs time kernels as they run unobserved.
    Uses the HIP runtime torch already loaded."""

    _hip = None
    FLAG = 0x20000000  # hipEventDisableSystemFence

    @classmethod
    def runtime(cls):
        import ctypes as C

        if cls._hip is None:
            with open("/proc/self/maps") as fh:
                path = next(ln.split()[-1] for ln in fh if "libamdhip64" in ln)
            cls._hip = C.CDLL(path)
        return cls._hip

    def __init__(self):
        import ctypes as C

        self.C = C
        self.h = C.c_void_p()
        if self.runtime().hipEventCreateWithFlags(C.byref(self.h), self.FLAG) != 0:
            raise RuntimeError("hipEventCreateWithFlags failed")

    def record(self, stream: int):
        if self.runtime().hipEventRecord(self.h, self.C.c_void_p(stream)) != 0:
            raise RuntimeError("hipEventRecord failed")

    def elapsed_ms(self, end: "FenceFreeEvent") -> float:
        ms = self.C.c_float()
        if self.runtime().hipEventElapsedTime(self.C.byref(ms), self.h, end.h) != 0:
            raise RuntimeError("hipEventElapsedTime failed")
        return ms.value

    def __del__(self):
        if self._hip is not None and self.h:
            self._hip.hipEventDestroy(self.h)


def held_frame_pass_times(s, sc):
    """Durations (ms) of every fused particle pass of one more frame of the workload.

    The frame's guard and sort run first (they read back to the host), then the stream
    is held by a GPU spin while the host enqueues the whole tick eagerly with a
    fence-free HIP event pair around every fused pass on its stream, so the events time
    the kernels back to back with no host launch gaps and no cache flushes of their own."""
    import torch

    from mpm_amd.core import native

    sim = s._sim
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    a.record()
    torch.cuda._sleep(1_000_000)
    b.record()
    torch.cuda.synchronize()
    spin_ms = a.elapsed_time(b)
    sc.command(s)
    s._update_grid_box(sc.dt, sc.substeps)
    s._sort_particles()
    saved = (s.guard_interval, s.sort_interval, sim.use_graph)
    s.guard_interval, s.sort_interval, sim.use_graph = 10 ** 9, 0, False
    events = []
    orig_pass = sim._pass
    fused_mode = native.PASS_G2P | native.PASS_STRESS | native.PASS_P2G

    def timed_pass(mode, dt, stream):
        if (mode & fused_mode) == fused_mode:
            e0, e1 = FenceFreeEvent(), FenceFreeEvent()
            e0.record(stream)
            orig_pass(mode, dt, stream)
            e1.record(stream)
            events.append((e0, e1))
        else:
            orig_pass(mode, dt, stream)

    sim._pass = timed_pass
    if spin_ms > 0:
        torch.cuda._sleep(int(1_000_000 * 150.0 / spin_ms))
    s.step(sc.dt, sc.substeps)
    torch.cuda.synchronize()
    sim._pass = orig_pass
    s.guard_interval, s.sort_interval, sim.use_graph = saved
    sc.advance()
    return [e0.elapsed_ms(e1) for e0, e1 in events]


def cdf_leg(args):
    """The pour with both glasses as CPIC (CDF) sheets along their cavities (the
    reference's --glass cdf pour): frames timed end to end (the moving source restamps
    every substep, so these ticks run eagerly), and the fused particle pass, which now
    carries the tag vote, ghost fill and blocked deposits, timed like the headline's."""
    import numpy as np
    import torch

    from mpm_amd import Solver
    from mpm_amd.scenes import PourScene

    sc = PourScene(glass="cdf")
    s = sc.build(Solver(grid=sc.grid, device="cuda:0", sort_interval=args.sort_interval,
                        sort_key=args.sort_key))
    n = s.n_particles
    frames = max(1, args.cdf_frames)
    for _ in range(args.warmup):
        sc.frame(s)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(frames):
        sc.frame(s)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    if not np.isfinite(s.x()).all():
        raise RuntimeError("non-finite particle state after the timed CDF frames")
    durations = held_frame_pass_times(s, sc)
    return {"value": round(n * frames * sc.substeps / el, 1), "unit": "particle-substeps/s",
            "ms_per_frame": round(el / frames * 1e3, 2), "frames": frames, "particles": n,
            "fused_pass_avg_us": round(float(np.mean(durations)) * 1e3, 2),
            "stamps": s._sim.cdf.stamp_count,
            "reference": "GH200 CDF pour: fused pass 2.93 ms/substep, ~2 s/frame "
                         "(docs/performance.md:99-105)"}


def float_leg(args):
    """The headline pour with Solver(deterministic=False): the float scatter (float atomics
    into the grid, the reference's accumulation, from per-workgroup fixed-point LDS sums)
    instead of the default 64-bit fixed-point integer one.
    Frame time and fused-pass time for comparison with the headline."""
    import numpy as np
    import torch

    from mpm_amd import Solver
    from mpm_amd.scenes import PourScene

    sc = PourScene()
    s = sc.build(Solver(grid=sc.grid, device="cuda:0", sort_interval=args.sort_interval,
                        sort_key=args.sort_key, deterministic=False))
    n = s.n_particles
    frames = max(1, args.float_frames)
    for _ in range(args.warmup):
        sc.frame(s)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(frames):
        sc.frame(s)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    if not np.isfinite(s.x()).all():
        raise RuntimeError("non-finite particle state after the float-atomic frames")
    durations = held_frame_pass_times(s, sc)
    return {"value": round(n * frames * sc.substeps / el, 1), "unit": "particle-substeps/s",
            "ms_per_frame": round(el / frames * 1e3, 2), "frames": frames + args.warmup,
            "fused_pass_avg_us": round(float(np.mean(durations)) * 1e3, 2)}


def _pour_frames(args):
    """The headline workload up to the end of the timed frames: (solver, scene, elapsed
    seconds of the timed frames, final-position digest)."""
    import hashlib

    import numpy as np
    import torch

    from mpm_amd import Solver
    from mpm_amd.scenes import PourScene

    sc = PourScene()
    # a sort at the start of every frame keeps each chunk's stencil box inside the LDS
    # arena of the particle pass; the cell key lets same-cell lanes sum before the atomics
    s = sc.build(Solver(grid=sc.grid, device="cuda:0", sort_interval=args.sort_interval,
                        sort_key=args.sort_key))
    s._sim.use_graph = not args.no_graph
    for _ in range(args.warmup):
        sc.frame(s)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        sc.frame(s)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    x_end = s.x()
    if not np.isfinite(x_end).all():
        raise RuntimeError("non-finite particle state after the timed frames")
    # the default scatter is bit-reproducible: this digest repeats from run to run
    digest = hashlib.sha256(np.ascontiguousarray(x_end).tobytes()).hexdigest()[:16]
    return s, sc, el, digest


def _active_nodes(s, sc):
    """Nodes receiving mass in one scatter of the current state (the default scatter
    accumulates into the fixed-point buffer, the float one into node_in)."""
    from mpm_amd.core import native

    sim = s._sim
    acc = sim.node_fix if sim.deterministic else sim.node_in
    acc.zero_()
    sim._pass(native.PASS_STRESS | native.PASS_P2G, sc.dt, sim.stream())
    active = int((acc.view(-1, 4)[:, 3] > 0).sum().item())
    acc.zero_()
    return active


def traffic_child(args):
    """Run under rocprofv3 by measured_traffic: the copy calibration, then the same pour
    as pour_leg up to and through its held frame; prints the final-position digest."""
    copy_calibration()
    s, sc, _, digest = _pour_frames(args)
    _active_nodes(s, sc)
    held_frame_pass_times(s, sc)
    print(json.dumps({"digest": digest}))


def pour_leg(args):
    """The headline: fused ticks of the 192^3 pour, frames timed end to end; the fused
    particle pass timed with HIP events on its stream; its HBM bytes and atomics counted
    by rocprofv3 on the same state."""
    import numpy as np
    import torch

    s, sc, el, digest = _pour_frames(args)
    n = s.n_particles
    substeps = args.steps * sc.substeps
    value = n * substeps / el
    active = _active_nodes(s, sc)
    durations = held_frame_pass_times(s, sc)
    avg_s = float(np.mean(durations)) / 1e3
    tiers = {}
    for name, (bp, bn) in TIERS.items():
        b = bp * n + bn * active
        tiers[name] = {"bytes_per_particle": bp, "bytes_per_active_node": bn, "algorithmic_bytes": int(b),
                       "achieved_GBs": round(b / avg_s / 1e9, 2), "frac": round(b / avg_s / 1e9 / HBM_PEAK_GBS, 4)}
    del s
    torch.cuda.empty_cache()
    traffic, why = (None, "--no-traffic") if args.no_traffic else measured_traffic(args, len(durations), avg_s, digest)
    tb = tiers["B"]
    roof = {"bound": "hbm", "achieved": tb["achieved_GBs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": tb["frac"], "traffic": None if traffic is None else traffic["bytes"],
            "kernel": "k_particle_pass (fused g2p+stress+p2g, fixed-point scatter)",
            "avg_launch_us": round(avg_s * 1e6, 2), "launches": len(durations),
            "algorithmic_bytes_per_launch": tb["algorithmic_bytes"], "active_nodes": active, "tier": "B",
            "tiers": tiers, "measured": traffic if traffic is not None else {"error": why}}
    return {"value": round(value, 1), "ms_per_step": round(el / args.steps * 1e3, 3), "particles": n,
            "substeps_per_step": sc.substeps, "dt": sc.dt, "roofline": roof, "x_sha256_16": digest}


def main():
    args = parse()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.traffic_child:
        traffic_child(args)
        return
    legs = set(args.legs.split(","))
    if args.no_batch:
        legs.discard("batch")
    if args.no_adjoint:
        legs.discard("adjoint")
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                                   "(multi-GPU sharding is a design only)"}))
        sys.exit(2)
    import torch

    pour = pour_leg(args) if "pour" in legs else {}
    torch.cuda.empty_cache()
    cpu = None if (args.no_cpu or "pour" not in legs) else cpu_baseline(args.cpu_substeps)
    adjoint = batch = None
    if "adjoint" in legs:
        try:
            adjoint = adjoint_leg(args)
        except Exception as e:  # reported in the line; the headline value stands
            adjoint = {"error": f"{type(e).__name__}: {e}"}
        torch.cuda.empty_cache()
    if "batch" in legs:
        try:
            batch = batch_leg(args)
        except Exception as e:  # reported in the line; the headline value stands
            batch = {"error": f"{type(e).__name__}: {e}"}
        torch.cuda.empty_cache()
    flt = None
    if "float" in legs:
        try:
            flt = float_leg(args)
        except Exception as e:  # reported in the line; the headline value stands
            flt = {"error": f"{type(e).__name__}: {e}"}
        torch.cuda.empty_cache()
    cdf = None
    if "cdf" in legs:
        try:
            cdf = cdf_leg(args)
        except Exception as e:  # reported in the line; the headline value stands
            cdf = {"error": f"{type(e).__name__}: {e}"}
    out = {"metric": "particle-substeps/s at 340k pour", "value": pour.get("value"),
           "unit": "particle-substeps/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": pour.get("ms_per_step"),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
           "data": "synthetic",
           "config": {"workload": "honey pour 192^3, eta 10 Pa s, 2 revolved-SDF glasses, "
                                  "source tilting 1 rad/s", "particles": pour.get("particles"),
                      "substeps_per_step": pour.get("substeps_per_step"), "dt": pour.get("dt"),
                      "grid": 192, "graph": not args.no_graph, "sort_interval": args.sort_interval,
                      "sort_key": args.sort_key, "scatter": "fixed-point (deterministic)",
                      "x_sha256_16": pour.get("x_sha256_16")},
           "roofline": pour.get("roofline"), "cpu_baseline": cpu, "batched": batch, "adjoint": adjoint,
           "cdf_pour": cdf, "float_atomics": flt}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
