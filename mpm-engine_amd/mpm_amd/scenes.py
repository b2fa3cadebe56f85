"""Scene builders shared by tests and the bench (reference: src/warpmpm/scenes.py)."""
from __future__ import annotations

import numpy as np

from .core.solver import GridConfig
from .materials import Material, newtonian


def dough(eta: float = 40.0, tau_y: float = 200.0) -> Material:
    """Bingham dough: Newtonian with a yield stress."""
    return newtonian(eta=eta, density=1000.0).with_yield(tau_y)


def block(grid: GridConfig, size=(0.12, 0.05, 0.06), center=None, ppc: int = 2, seed: int = 0):
    """Jittered particle box resting on the floor (3 cells up). Returns (pos, vol, floor)."""
    dx = grid.dx
    floor = 3 * dx
    h = dx / ppc
    cx = grid.grid_lim * 0.5 if center is None else center[0]
    cy = grid.grid_lim * 0.5 if center is None else center[1]
    sx, sy, sz = size
    axes = (np.arange(cx - 0.5 * sx + 0.5 * h, cx + 0.5 * sx, h),
            np.arange(cy - 0.5 * sy + 0.5 * h, cy + 0.5 * sy, h),
            np.arange(floor + 0.5 * h, floor + sz, h))
    g = np.stack(np.meshgrid(*axes, indexing="ij"), axis=-1).reshape(-1, 3)
    g = g + np.random.default_rng(seed).uniform(-0.25 * h, 0.25 * h, size=g.shape)
    return g.astype(np.float32), np.full(len(g), h ** 3, dtype=np.float32), floor


def lattice_box(lo, hi, h: float, seed: int = 0):
    """Jittered lattice (spacing h, jitter +-h/4) filling the box [lo, hi)."""
    axes = [np.arange(a + 0.5 * h, b, h) for a, b in zip(lo, hi)]
    g = np.stack(np.meshgrid(*axes, indexing="ij"), axis=-1).reshape(-1, 3)
    g = g + np.random.default_rng(seed).uniform(-0.25 * h, 0.25 * h, size=g.shape)
    return g.astype(np.float32), np.full(len(g), h ** 3, dtype=np.float32)


def cylinder(center, radius: float, height: float, h: float, seed: int = 0):
    """Jittered lattice filling an upright (z-axis) cylinder with its base at center."""
    cx, cy, z0 = center
    g, vol = lattice_box((cx - radius, cy - radius, z0), (cx + radius, cy + radius, z0 + height),
                         h, seed)
    keep = np.hypot(g[:, 0] - cx, g[:, 1] - cy) < radius
    return g[keep], vol[keep]


class PourScene:
    """Newtonian honey pour on a 192^3 grid over [0, 1]^3 (the production workload):
    honey (eta 10 Pa s, bulk 1e4 Pa, rho 1400) fills a cylinder of radius 0.08 at 8
    particles per cell inside a source glass whose wall is 2 cells thick; a static target
    glass stands beside it. The source tilts at omega_tilt about a pivot on its rim. One
    frame is 1/60 s in `substeps` substeps (dt = 1 / (60 * substeps)).

    glass="cdf" replaces the two revolved-SDF glasses by CPIC sheets along their cavities
    (the reference's --glass cdf pour, examples/pour_franka.py:249-290)."""

    def __init__(self, n_grid: int = 192, height: float = 0.30, radius: float = 0.08,
                 substeps: int = 432, omega_tilt: float = 1.0, seed: int = 0, glass: str = "cup"):
        from .colliders.glass import GlassProfile

        if glass not in ("cup", "cdf"):
            raise ValueError("glass must be 'cup' or 'cdf'")
        self.glass = glass
        self.grid = GridConfig(n_grid=n_grid, grid_lim=1.0)
        dx = self.grid.dx
        self.substeps = substeps
        self.dt = 1.0 / (60.0 * substeps)
        inner = radius + 2 * dx
        self.profile = GlassProfile(outer_radius=inner + 2 * dx, inner_radius=inner,
                                    height=height + 0.08, base_thickness=0.03, fillet_radius=0.01)
        p = self.profile
        self.source_center = np.array([0.32, 0.5, 0.06 + p.half_height])
        self.target_center = np.array([0.72, 0.5, 0.06 + p.half_height])
        floor = self.source_center[2] + p.inner_floor_z
        self.pos, self.vol = cylinder((0.32, 0.5, floor + dx), radius, height, dx / 2, seed)
        # rim pivot on the +x side of the source glass; tilt about +y
        self.pivot = self.source_center + np.array([p.outer_radius, 0.0, p.half_height])
        self.omega = np.array([0.0, omega_tilt, 0.0])
        self.center = self.source_center.astype(np.float64).copy()
        self.quat = np.array([1.0, 0.0, 0.0, 0.0])

    def material(self):
        return newtonian(eta=10.0, density=1400.0, bulk_modulus=1.0e4)

    def build(self, solver):
        solver.load_particles(self.pos, self.vol)
        solver.set_material(self.material())
        solver.add_plane((0, 0, 3 * self.grid.dx), (0, 0, 1), "separate", friction=0.3)
        solver.add_domain_walls()
        if self.glass == "cdf":
            from .colliders.glass import glass_cdf

            cdf = glass_cdf(self.profile, self.grid.dx)
            self.src = solver.add_cdf_collider(cdf, self.source_center, friction=0.05)
            self.dst = solver.add_cdf_collider(cdf, self.target_center, friction=0.05)
        else:
            self.src = solver.add_cup(self.profile, self.source_center)
            self.dst = solver.add_cup(self.profile, self.target_center)
        return solver

    def command(self, solver):
        """Command the source glass for the coming frame: its start-of-frame pose and the
        rigid velocity of the tilt about the pivot."""
        vel = np.cross(self.omega, self.center - self.pivot)
        if self.glass == "cdf":
            w, x, y, z = (float(q) for q in self.quat)
            solver.set_cdf_pose(self.src, center=self.center, quat=(x, y, z, w), velocity=vel,
                                omega=self.omega)
        else:
            solver.set_cup(self.src, center=self.center, quat=self.quat, velocity=vel, omega=self.omega)

    def advance(self):
        """Exact rigid motion of the commanded pose over one frame."""
        from .colliders.glass import quat_from_axis_angle, quat_mul, quat_to_mat

        ang = float(np.linalg.norm(self.omega)) * self.dt * self.substeps
        if ang > 0:
            ax = self.omega / np.linalg.norm(self.omega)
            q = quat_from_axis_angle(ax, ang)
            R = quat_to_mat(q)
            self.center = self.pivot + R @ (self.center - self.pivot)
            self.quat = quat_mul(q, self.quat)

    def frame(self, solver):
        """Command the source glass, advance one frame, move the commanded pose."""
        self.command(solver)
        solver.step(self.dt, self.substeps)
        self.advance()


def herschel_bulkley_dough() -> Material:
    """Herschel-Bulkley dough (K = 50 Pa s^n, n = 0.5, rho 1000); the yield stress is set
    per particle."""
    return newtonian(eta=0.0, density=1000.0).with_yield(200.0).with_powerlaw(K=50.0, n=0.5)


class DoughPressBatch:
    """Independent dough-press scenes on a 64^3 grid over [0, 1]^3: a seeded
    Herschel-Bulkley slab per scene (particle count drawn uniformly from
    [n_min, n_max] at 8 particles per cell, seeded footprint aspect ratio and height),
    per-particle yield stress log-uniform in [50, 500] Pa, a sticky floor 3 cells up,
    domain walls, and a kinematic box plate over the slab moving down at a seeded speed
    in [0.02, 0.1] m/s. dt = 5e-5."""

    def __init__(self, n_scenes: int = 256, n_grid: int = 64, n_min: int = 5000, n_max: int = 60000,
                 seed: int = 0, dt: float = 5e-5):
        self.grid = GridConfig(n_grid=n_grid, grid_lim=1.0)
        self.dt = dt
        dx = self.grid.dx
        h = dx / 2.0
        rng = np.random.default_rng(seed)
        self.floor = 3 * dx
        self.pos, self.vol, self.yield_stress = [], [], []
        self.box_center, self.box_half, self.box_speed = [], [], []
        side_max = 0.6
        for s in range(n_scenes):
            n_target = int(rng.integers(n_min, n_max + 1))
            aspect = float(np.exp(rng.uniform(np.log(0.5), np.log(2.0))))
            volume = n_target * h ** 3
            lz = max(float(rng.uniform(0.05, 0.12)), volume / (side_max * side_max / max(aspect, 1.0 / aspect)))
            area = volume / lz
            lx, ly = np.sqrt(area * aspect), np.sqrt(area / aspect)
            nx, ny, nz = (max(2, int(round(v / h))) for v in (lx, ly, lz))
            lx, ly, lz = nx * h, ny * h, nz * h
            lo = (0.5 - 0.5 * lx, 0.5 - 0.5 * ly, self.floor)
            pos, vol = lattice_box(lo, (lo[0] + lx, lo[1] + ly, lo[2] + lz), h, seed=seed * 100003 + s)
            self.pos.append(pos)
            self.vol.append(vol)
            self.yield_stress.append(np.exp(rng.uniform(np.log(50.0), np.log(500.0), size=len(pos))).astype(np.float32))
            half = (0.5 * lx + 2 * dx, 0.5 * ly + 2 * dx, 2 * dx)
            self.box_half.append(half)
            self.box_center.append((0.5, 0.5, self.floor + lz + 0.5 * dx + half[2]))
            self.box_speed.append(float(rng.uniform(0.02, 0.1)))
        self.counts = np.array([len(p) for p in self.pos])

    def build_batch(self, solver):
        """Load every scene into a BatchSolver and register its colliders."""
        solver.load_scenes(self.pos, self.vol)
        solver.set_material(herschel_bulkley_dough())
        solver.set_yield_stress(self.yield_stress)
        solver.add_plane((0, 0, self.floor), (0, 0, 1), "sticky")
        solver.add_domain_walls()
        self.tool = [solver.add_box(s, self.box_center[s], self.box_half[s], (0.0, 0.0, -self.box_speed[s]))
                     for s in range(len(self.pos))]
        return solver

    def build_one(self, solver, s: int):
        """Scene s alone in a plain Solver (the sequential baseline)."""
        solver.load_particles(self.pos[s], self.vol[s])
        solver.set_material(herschel_bulkley_dough())
        solver._sim.yield_stress.copy_(solver._sim.torch.from_numpy(self.yield_stress[s]).to(solver._sim.device))
        solver.add_plane((0, 0, self.floor), (0, 0, 1), "sticky")
        solver.add_domain_walls()
        self.tool_one = solver.add_box(self.box_center[s], self.box_half[s], (0.0, 0.0, -self.box_speed[s]))
        return solver


class SqueezeScene:
    """Squeeze identification: a seeded jittered 0.2^3 block of von Mises material
    (E = 2e5 Pa, nu = 0.3, yield 1e3 Pa, rho 1000) resting on a sticky floor 3 cells up,
    between two box plates that close along x at `plate_speed` each. `per_side`^3
    particles (37^3 = 50,653 at about 24 per cell on 64^3); dt = 1e-4."""

    def __init__(self, n_grid: int = 64, per_side: int = 37, side: float = 0.2, dt: float = 1e-4,
                 plate_speed: float = 0.05, seed: int = 0):
        from .materials import vonmises

        self.grid = GridConfig(n_grid=n_grid, grid_lim=1.0)
        dx = self.grid.dx
        self.dt = dt
        h = side / per_side
        self.floor = 3 * dx
        lo = (0.5 - 0.5 * side, 0.5 - 0.5 * side, self.floor + 0.5 * dx)
        self.pos, self.vol = lattice_box(lo, (lo[0] + side, lo[1] + side, lo[2] + side), h, seed)
        self.material = vonmises(E=2.0e5, nu=0.3, yield_stress=1.0e3, density=1000.0)
        t = 2 * dx
        self.plate_half = (t, 0.5 * side + 2 * dx, 0.5 * side + 2 * dx)
        zc = lo[2] + 0.5 * side
        self.plate_centers = ((lo[0] - t, 0.5, zc), (lo[0] + side + t, 0.5, zc))
        self.plate_velocity = ((plate_speed, 0.0, 0.0), (-plate_speed, 0.0, 0.0))

    def build(self, solver):
        solver.load_particles(self.pos, self.vol)
        solver.set_material(self.material)
        solver.add_plane((0, 0, self.floor), (0, 0, 1), "sticky")
        self.plates = [solver.add_box(c, self.plate_half, v)
                       for c, v in zip(self.plate_centers, self.plate_velocity)]
        return solver


class DoughSandPress:
    """Dough press beside a sand heap on a 128^3 grid over [0, 1]^3 (sparse blocks, block
    sort every `sort_interval` ticks): a Herschel-Bulkley dough slab 0.3 x 0.3 x 0.1
    (yield 200 Pa, K 50, n 0.5, rho 1000) and a mu(I) sand heap of seeded random height on
    a 0.3 x 0.3 footprint (mean height 0.085, mu_s 0.38, mu_2 = mu_s + 0.26, I0 0.28, d 1 mm, rho 1550), both
    at `ppc` particles per cell (16 gives ~300k each), a sticky floor 3 cells up, domain
    walls, and a kinematic 0.12 x 0.12 box touching the dough top pressing down at 0.05 m/s until the
    stress-integral reaction under it reaches `f_stop` (30 N), then holding. dt = 5e-5;
    a tick is `substeps` substeps (force checked once per tick)."""

    def __init__(self, n_grid: int = 128, ppc: int = 16, seed: int = 0, substeps: int = 20,
                 f_stop: float = 30.0, speed: float = 0.05, gap_cells: float = 0.0):
        self.grid = GridConfig(n_grid=n_grid, grid_lim=1.0)
        dx = self.grid.dx
        self.dt = 5e-5
        self.substeps = substeps
        self.f_stop, self.speed = f_stop, speed
        h = dx / ppc ** (1.0 / 3.0)
        self.floor = 3 * dx
        fl = self.floor
        rng = np.random.default_rng(seed)
        dpos, dvol = lattice_box((0.10, 0.35, fl), (0.40, 0.65, fl + 0.10), h, seed)
        # heap: columns of the footprint lattice, height = smooth seeded bumps in [0.05, 0.15]
        cols = np.stack(np.meshgrid(np.arange(0.55 + 0.5 * h, 0.85, h), np.arange(0.35 + 0.5 * h, 0.65, h),
                                    indexing="ij"), axis=-1).reshape(-1, 2)
        k = rng.uniform(4.0, 12.0, size=(3, 2))
        ph = rng.uniform(0.0, 2.0 * np.pi, size=3)
        bump = sum(np.sin(k[i, 0] * cols[:, 0] + ph[i]) * np.sin(k[i, 1] * cols[:, 1]) for i in range(3)) / 3.0
        height = 0.085 + 0.04 * bump
        spos = []
        for j, (x, y) in enumerate(cols):
            zs = np.arange(fl + 0.5 * h, fl + height[j], h)
            spos.append(np.stack([np.full(len(zs), x), np.full(len(zs), y), zs], axis=-1))
        spos = np.concatenate(spos)
        spos = spos + rng.uniform(-0.25 * h, 0.25 * h, size=spos.shape)
        self.n_dough = len(dpos)
        self.pos = np.concatenate([dpos, spos.astype(np.float32)])
        self.vol = np.concatenate([dvol, np.full(len(spos), h ** 3, np.float32)])
        self.box_half = np.array([0.06, 0.06, 0.02])
        # the plate starts `gap_cells` above the slab top (0: touching)
        self.box_center = np.array([0.25, 0.50, fl + 0.10 + 0.02 + gap_cells * dx])
        self.stopped = False

    def materials(self):
        from .materials import granular
        dough_m = newtonian(eta=0.0, density=1000.0).with_yield(200.0).with_powerlaw(K=50.0, n=0.5)
        sand = granular(mu_s=0.38, delta_mu=0.26, I0=0.28, grain_diameter=1.0e-3, density=1550.0)
        return dough_m, sand

    def build(self, solver):
        dough_m, sand = self.materials()
        solver.load_particles(self.pos, self.vol)
        solver.set_material(dough_m)
        solver.set_material_range(self.n_dough, len(self.pos), sand)
        solver.add_plane((0, 0, self.floor), (0, 0, 1), "sticky")
        solver.add_domain_walls()
        self.box = solver.add_box(self.box_center, self.box_half, velocity=(0.0, 0.0, -self.speed))
        return solver

    def tick(self, solver, wrench=None):
        """One tick, then the force check: the stress-integral Fz under the box (from
        `wrench(center, half)`, default the solver's device readout) stops the box once it
        reaches f_stop. Returns Fz."""
        solver.step(self.dt, self.substeps)
        if not self.stopped:
            self.box_center = self.box_center + np.array([0.0, 0.0, -self.speed * self.dt * self.substeps])
        fn = wrench or solver.box_contact_wrench
        fz = float(fn(self.box_center, self.box_half)["Fz"])
        if not self.stopped and fz >= self.f_stop:
            self.stopped = True
            solver.set_box(self.box, velocity=(0.0, 0.0, 0.0))
        return fz
