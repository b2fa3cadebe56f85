"""Oracle orchestration: numpy restatement of warpmpm's simulator host logic.

TEST INFRASTRUCTURE ONLY (the checker). Only tests/, the smoke check and the bench's
cpu_baseline leg import this module; the product Solver never does.

`OracleSim` restates ``MPM_Simulator_WARP`` (src/warpmpm/kernels/mpm_solver_warp.py): the
split substep ``p2g2p`` (:1092-1281), the fused tick ``p2g2p_fused_tick`` (:1379-1546), the
per-collider BC launches and host pose integration ``_apply_grid_bc`` (:1283-1336), the
collider registry (:1880-2814), sparse blocks (:989-1026) and the permutation (:1548-1619).
`OracleSolver` restates the typed wrapper ``Solver`` (src/warpmpm/core/solver.py:57-642).
Device kernels are the C functions in mpm_oracle.c, called through ctypes on numpy
arrays laid out like the reference's Warp arrays (AoS vec3 / row-major mat33, dense grid
with z fastest). Host-side numerics follow the reference: poses integrate in float64 and
are stored as float32 (Warp struct fields are fp32), kernel `time` arguments are fp32.
"""
from __future__ import annotations

import ctypes as C
import os
import warnings

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
MAX_MATERIALS = 16
# material ids, mpm_solver_warp.py:14-19
MATERIAL_NAME_TO_ID = {
    "jelly": 0, "metal": 1, "sand": 2, "plasticine": 5, "fluid": 6, "stationary": 7,
    "rigid": 8, "mu_i_sand": 9, "newtonian": 10, "mu_i_phi": 11, "tabulated_viscous": 12,
    "tabulated_mu_i": 13,
}

f32 = np.float32
FP = C.POINTER(C.c_float)
IP = C.POINTER(C.c_int32)


class OModel(C.Structure):
    _fields_ = [
        ("dx", C.c_float), ("inv_dx", C.c_float),
        ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
        ("gravity", C.c_float * 3), ("rpic_damping", C.c_float),
        ("grid_v_damping_scale", C.c_float), ("update_cov_with_F", C.c_int32),
        ("periodic_x", C.c_int32),
        ("E", FP), ("nu", FP), ("bulk", FP), ("hardening", FP), ("xi", FP),
        ("plastic_viscosity", FP), ("softening", FP), ("alpha", FP),
        ("muI_mu_s", FP), ("muI_delta_mu", FP), ("muI_I0", FP), ("muI_d", FP),
        ("muI_rho_s", FP), ("muI_phi_init", FP), ("muI_phi_chi", FP),
        ("eta_table", FP), ("eta_table_smin", C.c_float), ("eta_table_smax", C.c_float),
        ("eta_table_n", C.c_int32),
        ("mu", FP), ("lam", FP), ("yield_stress", FP),
    ]


MAX_CDF = 4
_LANE3 = (C.c_float * 3) * MAX_CDF


class OCdf(C.Structure):
    _fields_ = [("n_cdf", C.c_int32), ("mask_off", C.c_int32), ("nb", C.c_int32 * 3),
                ("tag", IP), ("tag_prev", IP), ("d", FP), ("d_prev", FP), ("block", IP),
                ("ptag", IP),
                ("lane_center", _LANE3), ("lane_velocity", _LANE3), ("lane_omega", _LANE3),
                ("lane_center_prev", _LANE3), ("lane_velocity_prev", _LANE3),
                ("lane_omega_prev", _LANE3), ("lane_friction", C.c_float * MAX_CDF),
                ("lane_type", C.c_int32 * MAX_CDF), ("reaction_force", _LANE3),
                ("reaction_torque", _LANE3)]


class OCdfCollider(C.Structure):
    _fields_ = [("cdf_val", FP), ("cdf_valid", FP), ("res", C.c_int32),
                ("origin", C.c_float * 3), ("cell", C.c_float), ("center", C.c_float * 3),
                ("quat", C.c_float * 4), ("band", C.c_float), ("lane", C.c_int32),
                ("start_time", C.c_float), ("end_time", C.c_float)]


class OState(C.Structure):
    _fields_ = [
        ("n", C.c_int32),
        ("x", FP), ("v", FP), ("F", FP), ("F_trial", FP), ("C", FP), ("stress", FP),
        ("L", FP), ("R", FP), ("vol", FP), ("mass", FP), ("density", FP), ("Jp", FP),
        ("cov", FP), ("init_cov", FP), ("selection", IP), ("material", IP),
        ("grid_m", FP), ("grid_v_in", FP), ("grid_v_out", FP),
        ("cdf", C.POINTER(OCdf)), ("rigid_id", IP), ("x_ref", FP),
    ]


class OPlane(C.Structure):
    _fields_ = [("start_time", C.c_float), ("end_time", C.c_float),
                ("point", C.c_float * 3), ("normal", C.c_float * 3),
                ("surface_type", C.c_int32), ("friction", C.c_float)]


class OBox(C.Structure):
    _fields_ = [("start_time", C.c_float), ("end_time", C.c_float),
                ("point", C.c_float * 3), ("size", C.c_float * 3),
                ("velocity", C.c_float * 3), ("reset", C.c_int32), ("force", FP)]


class OCup(C.Structure):
    _fields_ = [("start_time", C.c_float), ("end_time", C.c_float),
                ("point", C.c_float * 3), ("rot", C.c_float * 9),
                ("velocity", C.c_float * 3), ("omega", C.c_float * 3),
                ("friction", C.c_float), ("outer_radius", C.c_float),
                ("inner_radius", C.c_float), ("half_height", C.c_float),
                ("inner_floor_z", C.c_float), ("fillet_radius", C.c_float),
                ("sticky_depth", C.c_float), ("contact_band", C.c_float),
                ("force", FP), ("torque", FP)]


class OWalls(C.Structure):
    _fields_ = [("start_time", C.c_float), ("end_time", C.c_float)]


class OSdf(C.Structure):
    _fields_ = [("sdf_val", FP), ("sdf_grad", FP), ("res", C.c_int32),
                ("origin", C.c_float * 3), ("cell", C.c_float), ("center", C.c_float * 3),
                ("quat", C.c_float * 4), ("velocity", C.c_float * 3),
                ("omega", C.c_float * 3), ("band", C.c_float), ("surface_type", C.c_int32),
                ("friction", C.c_float), ("start_time", C.c_float), ("end_time", C.c_float),
                ("force", FP), ("torque", FP)]


class OPointCloud(C.Structure):
    _fields_ = [("occupancy", IP), ("start_time", C.c_float), ("end_time", C.c_float)]


_LIB = None
_LIB_NAME = "libmpm_oracle.so"


def use_threaded_build(on: bool) -> None:
    """Switch to the OpenMP build (every host core, atomic scatter) for CPU-baseline timing,
    or back to the single-threaded checker build."""
    global _LIB, _LIB_NAME
    name = "libmpm_oracle_omp.so" if on else "libmpm_oracle.so"
    if name != _LIB_NAME:
        _LIB_NAME, _LIB = name, None


def _variant_path(name: str, flags) -> str:
    """The oracle compiled with extra flags into oracle/_mut/libmpm_oracle_<name>.so on
    first use (never the checker)."""
    import subprocess

    d = os.path.join(_HERE, "_mut")
    os.makedirs(d, exist_ok=True)
    path = os.path.join(d, f"libmpm_oracle_{name}.so")
    src = os.path.join(_HERE, "mpm_oracle.c")
    if not os.path.exists(path) or os.path.getmtime(path) < os.path.getmtime(src):
        tmp = f"{path}.{os.getpid()}.tmp"
        subprocess.run(["gcc", "-O2", "-fPIC", "-std=c11", "-fno-fast-math", *flags, "-shared", "-o", tmp, src,
                        "-lm"], check=True, capture_output=True)
        os.replace(tmp, path)
    return path


def mutant_path(k: int) -> str:
    """Path of the oracle built with planted mistake k (-DO_MUTANT=k, see MUT(k) in
    mpm_oracle.c). Tests load it to show that a pin fails on the mistake."""
    return _variant_path(f"mut{int(k)}", ["-ffp-contract=off", f"-DO_MUTANT={int(k)}"])


def contracted_path() -> str:
    """The oracle with fused multiply-adds (-mfma -ffp-contract=fast): the same arithmetic
    with other roundings in every substep, the kind of difference a GPU build has from the
    checker. Tests use it to measure a workload's conditioning; it is never the checker."""
    return _variant_path("fma2", ["-mfma", "-ffp-contract=fast"])


class mutant:
    """Context manager: the oracle functions run the build with planted mistake k (or,
    with k = "fma", the contracted build)."""

    def __init__(self, k):
        self.path = contracted_path() if k == "fma" else mutant_path(k)

    def __enter__(self):
        global _LIB, _LIB_NAME
        self.saved = (_LIB, _LIB_NAME)
        _LIB, _LIB_NAME = None, self.path
        return self

    def __exit__(self, *exc):
        global _LIB, _LIB_NAME
        _LIB, _LIB_NAME = self.saved
        return False


def lib():
    """Load (building on first use if needed) the oracle shared library."""
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, _LIB_NAME)
        if not os.path.exists(path):
            import subprocess
            subprocess.run(["make", "-C", _HERE], check=True, capture_output=True)
        L = C.CDLL(path)
        S, M = C.POINTER(OState), C.POINTER(OModel)
        L.o_svd3.argtypes = [FP, FP, FP, FP]
        L.o_compute_mu_lam.argtypes = [S, M]
        for nm in ("o_compute_stress", "o_p2g", "o_g2p", "o_g2p_stress_p2g"):
            getattr(L, nm).argtypes = [S, M, C.c_float]
        L.o_stress_particle.argtypes = [S, M, C.c_float, C.c_int]
        L.o_zero_grid.argtypes = [S, M, IP, IP, C.c_int]
        L.o_grid_normalize.argtypes = [S, M, C.c_float, IP, IP]
        L.o_grid_damping.argtypes = [S, M, C.c_float, IP, IP]
        for nm, st in (("o_collide_plane", OPlane), ("o_collide_box", OBox),
                       ("o_collide_cup", OCup), ("o_collide_walls", OWalls),
                       ("o_collide_sdf", OSdf), ("o_collide_pointcloud", OPointCloud)):
            getattr(L, nm).argtypes = [C.c_float, C.c_float, S, M, C.POINTER(st), IP, IP]
        L.o_blocks_mark.argtypes = [FP, C.c_int, C.c_float, IP, IP]
        L.o_blocks_dilate.argtypes = [IP, IP, IP]
        L.o_blocks_compact.argtypes = [IP, IP, IP, IP, IP, IP]
        L.o_blocks_grid_op.argtypes = [S, M, C.c_int, C.c_float, C.c_float, IP, C.c_int, IP]
        for nm in ("o_cdf_zero_tags", "o_cdf_copy_prev", "o_cdf_block_mask"):
            getattr(L, nm).argtypes = [S, M, IP, IP]
        L.o_cdf_stamp.argtypes = [C.c_float, S, M, C.POINTER(OCdfCollider), IP, IP]
        L.o_compute_R.argtypes = [S]
        L.o_compute_cov.argtypes = [S]
        L.o_rigid_gather.argtypes = [S, M, FP, C.c_int, FP, FP]
        L.o_rigid_integrate.argtypes = [FP] * 8 + [C.c_int, C.c_float]
        L.o_rigid_update.argtypes = [S, FP, FP, FP, FP, C.c_int]
        _LIB = L
    return _LIB


def fptr(a):
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(FP)


def iptr(a):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(IP)


def i3(t):
    return (C.c_int32 * 3)(*[int(v) for v in t])


def svd3(A):
    """Oracle 3x3 SVD (float32 in / out): A = U diag(s) V^T."""
    A = np.ascontiguousarray(A, dtype=f32).reshape(9)
    U = np.zeros(9, f32); s = np.zeros(3, f32); V = np.zeros(9, f32)
    lib().o_svd3(fptr(A), fptr(U), fptr(s), fptr(V))
    return U.reshape(3, 3), s, V.reshape(3, 3)


def quat_to_mat(q):
    """Rotation matrix of a wxyz quaternion (restates colliders/glass.py:24-34)."""
    w, x, y, z = (float(v) for v in q)
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
        [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
        [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]], dtype=np.float64)


def _omega_step_quat(q, omega, dt):
    """xyzw exp-map step, world omega left-multiplies (mpm_solver_warp.py:186-210)."""
    w = np.asarray(omega, dtype=float)
    speed = float(np.linalg.norm(w))
    if speed * dt < 1e-12:
        return q
    half = 0.5 * speed * dt
    ax = w / speed
    dq = np.array([ax[0] * np.sin(half), ax[1] * np.sin(half), ax[2] * np.sin(half),
                   np.cos(half)])
    x1, y1, z1, w1 = dq
    x2, y2, z2, w2 = q
    qn = np.array([w1 * x2 + x1 * w2 + y1 * z2 - z1 * y2,
                   w1 * y2 - x1 * z2 + y1 * w2 + z1 * x2,
                   w1 * z2 + x1 * y2 - y1 * x2 + z1 * w2,
                   w1 * w2 - x1 * x2 - y1 * y2 - z1 * z2])
    return qn / np.linalg.norm(qn)


def _fa(vals):
    return np.asarray(vals, dtype=np.float64).astype(f32)


class OracleSim:
    """Restatement of MPM_Simulator_WARP (mpm_solver_warp.py:213-504 allocations)."""

    def __init__(self, pos, vol, n_grid, grid_lim, cov=None):
        n = len(pos)
        self.n_particles = n
        self.n_grid = int(n_grid)
        # model dx/inv_dx are fp32 struct fields (:225-235)
        self.dx = float(f32(grid_lim / n_grid))
        self.inv_dx = float(f32(float(n_grid / grid_lim)))
        self.grid_lim = float(f32(grid_lim))
        self.x = np.ascontiguousarray(pos, dtype=f32).reshape(n, 3).copy()
        self.v = np.zeros((n, 3), f32)
        eye = np.tile(np.eye(3, dtype=f32), (n, 1, 1))
        self.F = eye.copy()
        self.F_trial = eye.copy()
        self.C = np.zeros((n, 3, 3), f32)
        self.stress = np.zeros((n, 3, 3), f32)
        self.L = np.zeros((n, 3, 3), f32)
        self.R = np.zeros((n, 3, 3), f32)
        self.vol = np.ascontiguousarray(vol, dtype=f32).copy()
        self.mass = np.zeros(n, f32)
        self.density = np.zeros(n, f32)
        self.Jp = np.zeros(n, f32)
        self.init_cov = np.zeros(6 * n, f32)
        self.cov = np.zeros(6 * n, f32)
        if cov is not None:
            self.init_cov[:] = np.asarray(cov, f32).reshape(-1)
        self.selection = np.zeros(n, np.int32)
        self.material = np.zeros(n, np.int32)
        # rigid bodies (mpm_solver_warp.py:366-371, 486-504)
        self.rigid_id = np.zeros(n, np.int32)
        self.x_ref = np.zeros((n, 3), f32)
        self.n_rigid_bodies = 0
        self.rigid_surface_colliders = []
        # particle modifiers (mpm_solver_warp.py:486-490): parameters and selection masks
        self.impulse_params = []
        self.particle_velocity_modifier_params = []
        g = self.n_grid
        self.grid_m = np.zeros((g, g, g), f32)
        self.grid_v_in = np.zeros((g, g, g, 3), f32)
        self.grid_v_out = np.zeros((g, g, g, 3), f32)
        # per-type arrays with the reference defaults (:238-293)
        z = lambda: np.zeros(MAX_MATERIALS, f32)  # noqa: E731
        self.E, self.nu, self.bulk, self.hardening = z(), z(), z(), z()
        self.xi, self.plastic_viscosity = z(), z()
        sin_phi = float(f32(np.sin(f32(25.0 / 180.0 * 3.14159265))))
        alpha0 = float(f32(np.sqrt(f32(2.0 / 3.0)) * f32(2.0) * f32(sin_phi) / f32(3.0 - sin_phi)))
        self.softening = np.full(MAX_MATERIALS, 0.1, f32)
        self.alpha = np.full(MAX_MATERIALS, alpha0, f32)
        self.muI_mu_s, self.muI_delta_mu = z(), z()
        self.muI_I0 = np.ones(MAX_MATERIALS, f32)
        self.muI_d = np.ones(MAX_MATERIALS, f32)
        self.muI_rho_s = np.ones(MAX_MATERIALS, f32)
        self.muI_phi_init = np.ones(MAX_MATERIALS, f32)
        self.muI_phi_chi = z()
        self.eta_table = np.zeros(2, f32)
        self.eta_table_smin, self.eta_table_smax, self.eta_table_n = -1.0, 2.0, 2
        self.mu = np.zeros(n, f32)
        self.lam = np.zeros(n, f32)
        self.yield_stress = np.zeros(n, f32)
        self.gravity = np.zeros(3, f32)
        self.rpic_damping = 0.0
        self.grid_v_damping_scale = 1.1
        self.update_cov_with_F = 0
        self.periodic_x = 0
        self.material_id = 0
        self.time = 0.0
        # collider registry: five parallel lists (:440-449)
        self.colliders = []      # (kind, ctypes param)
        self.modify_bc = []
        self.collider_aabbs = []
        self.collider_labels = []
        self._keep = []          # keep accumulator arrays alive
        self._bc_box_cache = {}
        self.restrict_bc = True
        self.restrict_grid = True
        self.grid_launch_box = None
        self._prev_grid_box = None
        self.sparse_ready = False
        self._blk = None
        self._revolved_rot = {}
        self._sdf_guard = {}
        # CPIC (CDF) colliders, mpm_solver_warp.py:389-436: grids allocated on first add
        self.particle_cdf_tag = np.zeros(n, np.int32)
        self.cdf = None          # OCdf (tag grids, lane arrays, reaction accumulators)
        self.cdf_params = []
        self.cdf_aabbs = []
        self._cdf_guard = {}
        self._cdf_lane_dirty = []
        self._cdf_lane_active = []
        self._cdf_lane_prev_box = []
        self._cdf_lane_prev_stale = False
        self._cdf_tag_prev_stale_box = None
        self._cdf_stamp_count = 0
        self._cdf_force_stamp = False
        self._cdf_lag_prev = False

    # ------------------------------------------------------------------ ctypes views
    def _state(self):
        s = OState()
        s.n = self.n_particles
        for nm in ("x", "v", "F", "F_trial", "C", "stress", "L", "R", "vol", "mass",
                   "density", "Jp", "cov", "init_cov", "grid_m", "grid_v_in", "grid_v_out"):
            setattr(s, nm, fptr(getattr(self, nm)))
        s.selection = iptr(self.selection)
        s.material = iptr(self.material)
        s.rigid_id = iptr(self.rigid_id)
        s.x_ref = fptr(self.x_ref)
        if self.cdf is not None and self.cdf.n_cdf > 0:
            self.cdf.ptag = iptr(self.particle_cdf_tag)
            s.cdf = C.pointer(self.cdf)
        self._sref = s
        return C.byref(s)

    def _model(self):
        m = OModel()
        m.dx, m.inv_dx = self.dx, self.inv_dx
        m.nx = m.ny = m.nz = self.n_grid
        m.gravity[:] = [float(v) for v in self.gravity]
        m.rpic_damping = self.rpic_damping
        m.grid_v_damping_scale = self.grid_v_damping_scale
        m.update_cov_with_F = int(self.update_cov_with_F)
        m.periodic_x = int(self.periodic_x)
        for nm in ("E", "nu", "bulk", "hardening", "xi", "plastic_viscosity", "softening",
                   "alpha", "muI_mu_s", "muI_delta_mu", "muI_I0", "muI_d", "muI_rho_s",
                   "muI_phi_init", "muI_phi_chi", "eta_table", "mu", "lam", "yield_stress"):
            setattr(m, nm, fptr(getattr(self, nm)))
        m.eta_table_smin = self.eta_table_smin
        m.eta_table_smax = self.eta_table_smax
        m.eta_table_n = self.eta_table_n
        self._mref = m
        return C.byref(m)

    # ------------------------------------------------------------------ parameters
    def _material_name_to_id(self, name):
        if isinstance(name, (int, np.integer)):
            return int(name)
        if name not in MATERIAL_NAME_TO_ID:
            raise TypeError(f"Undefined material type: {name}")
        return MATERIAL_NAME_TO_ID[name]

    def set_parameters_dict(self, kw):
        """Restates set_parameters_dict, mpm_solver_warp.py:634-800."""
        if "material" in kw:
            self.material_id = self._material_name_to_id(kw["material"])
            self.material[:] = self.material_id
        # grid arrays are reallocated (zeroed) by every call (:659-673)
        self.grid_m[:] = 0
        self.grid_v_in[:] = 0
        self.grid_v_out[:] = 0
        m = self.material_id
        for key, arr in (("E", self.E), ("nu", self.nu), ("bulk_modulus", self.bulk)):
            if key in kw:
                arr[m] = kw[key]
        if "yield_stress" in kw:
            self.yield_stress[:] = kw["yield_stress"]
        if "hardening" in kw:
            self.hardening[m] = float(kw["hardening"])
        if "xi" in kw:
            self.xi[m] = kw["xi"]
        if "friction_angle" in kw:
            sin_phi = float(np.sin(kw["friction_angle"] / 180.0 * 3.14159265))
            self.alpha[m] = np.sqrt(2.0 / 3.0) * 2.0 * sin_phi / (3.0 - sin_phi)
        if "plastic_viscosity" in kw:
            self.plastic_viscosity[m] = kw["plastic_viscosity"]
        if "softening" in kw:
            self.softening[m] = kw["softening"]
        if "eta_table" in kw:
            self.eta_table = np.asarray(kw["eta_table"], dtype=f32).reshape(-1).copy()
            self.eta_table_n = int(self.eta_table.shape[0])
        if "eta_table_smin" in kw:
            self.eta_table_smin = float(kw["eta_table_smin"])
        if "eta_table_smax" in kw:
            self.eta_table_smax = float(kw["eta_table_smax"])
        for key, arr in (("mu_s", self.muI_mu_s), ("delta_mu", self.muI_delta_mu),
                         ("I0", self.muI_I0), ("grain_diameter", self.muI_d),
                         ("grain_density", self.muI_rho_s), ("phi_init", self.muI_phi_init),
                         ("phi_chi", self.muI_phi_chi)):
            if key in kw:
                arr[m] = float(kw[key])
        if "g" in kw:
            self.gravity[:] = _fa(kw["g"])
        if "density" in kw:
            self.density[:] = kw["density"]
            self.mass[:] = self.density * self.vol
        if "E" in kw or "nu" in kw:
            lib().o_compute_mu_lam(self._state(), self._model())
        if "rpic_damping" in kw:
            self.rpic_damping = kw["rpic_damping"]
        if "grid_v_damping_scale" in kw:
            self.grid_v_damping_scale = kw["grid_v_damping_scale"]

    def finalize_mu_lam(self):
        lib().o_compute_mu_lam(self._state(), self._model())

    def set_parameters_for_particles(self, a, b, kw):
        """Restates set_parameters_for_particles, mpm_solver_warp.py:1700-1783."""
        m = None
        if "material" in kw:
            m = self._material_name_to_id(kw["material"])
            self.material[a:b] = m
        if m is None:
            m = self.material_id
        for key, arr in (("E", self.E), ("nu", self.nu), ("bulk_modulus", self.bulk)):
            if key in kw:
                arr[m] = kw[key]
        if "hardening" in kw:
            self.hardening[m] = float(kw["hardening"])
        if "xi" in kw:
            self.xi[m] = kw["xi"]
        if "friction_angle" in kw:
            sin_phi = float(np.sin(kw["friction_angle"] / 180.0 * 3.14159265))
            self.alpha[m] = np.sqrt(2.0 / 3.0) * 2.0 * sin_phi / (3.0 - sin_phi)
        if "plastic_viscosity" in kw:
            self.plastic_viscosity[m] = kw["plastic_viscosity"]
        if "softening" in kw:
            self.softening[m] = kw["softening"]
        for key, arr in (("mu_s", self.muI_mu_s), ("delta_mu", self.muI_delta_mu),
                         ("I0", self.muI_I0), ("grain_diameter", self.muI_d),
                         ("grain_density", self.muI_rho_s), ("phi_init", self.muI_phi_init),
                         ("phi_chi", self.muI_phi_chi)):
            if key in kw:
                arr[m] = float(kw[key])
        if "density" in kw:
            self.density[a:b] = kw["density"]
            self.mass[a:b] = self.density[a:b] * self.vol[a:b]
        if "yield_stress" in kw:
            self.yield_stress[a:b] = kw["yield_stress"]
        if "E" in kw or "nu" in kw:
            E, nu = float(self.E[m]), float(self.nu[m])
            self.mu[a:b] = E / (2.0 * (1.0 + nu))
            self.lam[a:b] = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu))
        if "obj_id" in kw:  # rigid body id (mpm_solver_warp.py:1755-1758)
            self.rigid_id[a:b] = int(kw["obj_id"])

    # ------------------------------------------------------------------ rigid bodies
    def finalize_rigid_bodies(self):
        """mpm_solver_warp.py:834-873: per body (material 8, grouped by obj_id) the mass,
        the centre of mass, the inverse body-frame inertia (float64 sum, float32 store)
        and the body-frame reference position of each particle; orientation I."""
        rigid = self.material == 8
        if not np.any(rigid):
            return
        nb = int(self.rigid_id[rigid].max()) + 1
        self.n_rigid_bodies = nb
        self.rigid_x_cm = np.zeros((nb, 3), f32)
        self.rigid_v_cm = np.zeros((nb, 3), f32)
        self.rigid_omega = np.zeros((nb, 3), f32)
        self.rigid_orientation = np.tile(np.eye(3, dtype=f32), (nb, 1, 1))
        self.rigid_mass = np.zeros(nb, f32)
        self.rigid_inv_inertia_body = np.zeros((nb, 3, 3), f32)
        for b in range(nb):
            idx = np.where(rigid & (self.rigid_id == b))[0]
            m_b = self.mass[idx]
            self.rigid_mass[b] = float(m_b.sum())
            self.rigid_x_cm[b] = (self.x[idx] * m_b[:, None]).sum(axis=0) / self.rigid_mass[b]
        for b in range(nb):
            idx = np.where(rigid & (self.rigid_id == b))[0]
            r = self.x[idx] - self.rigid_x_cm[b]
            self.x_ref[idx] = r
            inertia = np.zeros((3, 3), np.float64)
            for ri, mi in zip(r, self.mass[idx]):
                inertia += mi * (np.dot(ri, ri) * np.eye(3) - np.outer(ri, ri))
            self.rigid_inv_inertia_body[b] = np.linalg.inv(inertia).astype(f32)

    def _apply_rigid_restitution(self):
        """mpm_solver_warp.py:874-962 (host numpy, float64): per body and surface, push the
        body out by the deepest penetration, then the normal impulse -(1 + e) v_n / denom at
        the deepest particle and a Coulomb-capped friction impulse."""
        if not self.rigid_surface_colliders or self.n_rigid_bodies == 0:
            return
        for b in range(self.n_rigid_bodies):
            idx = np.where((self.material == 8) & (self.rigid_id == b))[0]
            if len(idx) == 0:
                continue
            R = self.rigid_orientation[b]
            x_cm = self.rigid_x_cm[b]  # a view: pushes below show through it
            M = float(self.rigid_mass[b])
            I_world_inv = R @ self.rigid_inv_inertia_body[b] @ R.T
            x_world = x_cm + (R @ self.x_ref[idx].T).T
            for surf in self.rigid_surface_colliders:
                if self.time < surf["start_time"] or self.time >= surf["end_time"]:
                    continue
                pt = np.array(surf["point"], dtype=np.float64)
                n = np.array(surf["normal"], dtype=np.float64)
                e, mu = float(surf["restitution"]), float(surf["friction"])
                dists = (x_world - pt) @ n
                min_i = int(np.argmin(dists))
                if dists[min_i] >= 0.0:
                    continue
                push = (-float(dists[min_i])) * n
                self.rigid_x_cm[b] = self.rigid_x_cm[b] + push
                x_world = x_world + push
                r = x_world[min_i].astype(np.float64) - x_cm.astype(np.float64)
                v_contact = self.rigid_v_cm[b].astype(np.float64) + np.cross(
                    self.rigid_omega[b].astype(np.float64), r)
                v_n = float(np.dot(v_contact, n))
                if v_n >= 0.0:
                    continue
                r_cross_n = np.cross(r, n)
                denom = 1.0 / M + np.dot(np.cross(I_world_inv @ r_cross_n, r), n)
                J_n = -(1.0 + e) * v_n / denom
                self.rigid_v_cm[b] = self.rigid_v_cm[b].astype(np.float64) + (J_n / M) * n
                self.rigid_omega[b] = self.rigid_omega[b].astype(np.float64) + I_world_inv @ (J_n * r_cross_n)
                if mu > 0.0:
                    v_t_vec = v_contact - v_n * n
                    v_t_mag = float(np.linalg.norm(v_t_vec))
                    if v_t_mag > 1e-10:
                        t_hat = v_t_vec / v_t_mag
                        r_cross_t = np.cross(r, t_hat)
                        denom_t = 1.0 / M + np.dot(np.cross(I_world_inv @ r_cross_t, r), t_hat)
                        J_t = min(v_t_mag / denom_t, mu * J_n)
                        self.rigid_v_cm[b] = self.rigid_v_cm[b] - (J_t / M) * t_hat
                        self.rigid_omega[b] = self.rigid_omega[b] - I_world_inv @ (J_t * r_cross_t)

    # ------------------------------------------------------------ particle modifiers
    def _box_mask(self, point, size):
        off = self.x - np.asarray(point, f32)
        return np.all(np.abs(off) < np.asarray(size, f32), axis=1).astype(np.int32)

    def add_impulse_on_particles(self, force, dt, point=(1, 1, 1), size=(1, 1, 1), num_dt=1, start_time=0.0):
        """mpm_solver_warp.py:2385-2427 with selection_add_impulse_on_particles
        (mpm_utils.py:1320-1331)."""
        self.impulse_params.append({
            "start_time": start_time, "end_time": start_time + dt * num_dt,
            "force": np.asarray(force, f32), "mask": self._box_mask(point, size)})

    def enforce_particle_velocity_translation(self, point, size, velocity, start_time, end_time):
        """mpm_solver_warp.py:2431-2469 with selection_enforce_particle_velocity_translation
        (mpm_utils.py:1334-1345)."""
        self.particle_velocity_modifier_params.append({
            "kind": "translation", "start_time": start_time, "end_time": end_time,
            "velocity": np.asarray(velocity, f32), "mask": self._box_mask(point, size)})

    def enforce_particle_velocity_rotation(self, point, normal, half_height_and_radius, rotation_scale,
                                           translation_scale, start_time, end_time):
        """mpm_solver_warp.py:2475-2539 with selection_enforce_particle_velocity_cylinder
        (mpm_utils.py:1348-1363)."""
        n = np.asarray(normal, np.float64)
        n = ((1.0 / np.sqrt((n * n).sum())) * n).astype(f32)
        h1 = np.ones(3, f32)
        if abs(float(np.dot(n, h1))) < 0.01:
            h1 = np.array([0.72, 0.37, -0.67], f32)
        h1 = (h1 - np.dot(h1, n) * n).astype(f32)
        h1 = (h1 * (f32(1.0) / np.sqrt(np.dot(h1, h1)))).astype(f32)
        h2 = np.cross(h1, n).astype(f32)
        pt = np.asarray(point, f32)
        off = self.x - pt
        ax = off @ n
        hd = np.linalg.norm(off - ax[:, None] * n, axis=1)
        mask = ((np.abs(ax) < f32(half_height_and_radius[0])) & (hd < f32(half_height_and_radius[1]))).astype(np.int32)
        self.particle_velocity_modifier_params.append({
            "kind": "rotation", "start_time": start_time, "end_time": end_time, "point": pt, "normal": n,
            "axis1": h1, "axis2": h2, "rotation_scale": f32(rotation_scale),
            "translation_scale": f32(translation_scale), "mask": mask})

    def _apply_particle_modifiers(self, dt):
        """The split substep's pre_p2g_operations, then particle_velocity_modifiers
        (mpm_solver_warp.py:1181-1196), each on its selected particles while active."""
        t = self.time
        for p in self.impulse_params:
            if p["start_time"] <= t < p["end_time"]:
                sel = p["mask"] == 1
                m = self.mass[sel][:, None]
                self.v[sel] = self.v[sel] + (p["force"][None, :] / m) * f32(dt)
        for p in self.particle_velocity_modifier_params:
            if not (p["start_time"] <= t < p["end_time"]):
                continue
            sel = p["mask"] == 1
            if p["kind"] == "translation":
                self.v[sel] = p["velocity"]
                continue
            n, a1, a2 = p["normal"], p["axis1"], p["axis2"]
            off = self.x[sel] - p["point"]
            hd = np.linalg.norm(off - (off @ n)[:, None] * n, axis=1).astype(f32)
            th = np.arccos(np.clip((off @ a1) / hd, -1.0, 1.0)).astype(f32)
            th = np.where(off @ a2 > 0, th, -th)
            s1 = (-hd * np.sin(th) * p["rotation_scale"]).astype(f32)
            s2 = (hd * np.cos(th) * p["rotation_scale"]).astype(f32)
            self.v[sel] = (s1[:, None] * a1 + s2[:, None] * a2 + p["translation_scale"] * n).astype(f32)

    def _p2g2p_tail(self, dt):
        """mpm_solver_warp.py:1338-1377: the rigid step, then time += dt."""
        if self.n_rigid_bodies > 0:
            L = lib()
            nb = self.n_rigid_bodies
            lin = np.zeros((nb, 3), f32)
            ang = np.zeros((nb, 3), f32)
            L.o_rigid_gather(self._state(), self._model(), fptr(self.rigid_x_cm), nb, fptr(lin), fptr(ang))
            L.o_rigid_integrate(fptr(self.rigid_x_cm), fptr(self.rigid_v_cm), fptr(self.rigid_omega),
                                fptr(self.rigid_orientation), fptr(self.rigid_mass),
                                fptr(self.rigid_inv_inertia_body), fptr(lin), fptr(ang), nb, dt)
            self._apply_rigid_restitution()
            L.o_rigid_update(self._state(), fptr(self.rigid_x_cm), fptr(self.rigid_v_cm),
                             fptr(self.rigid_omega), fptr(self.rigid_orientation), nb)
        self.time = self.time + dt

    # ------------------------------------------------------------------ grid boxes
    def _grid_box(self, lo_w, hi_w, halo=1):
        """mpm_solver_warp.py:1868-1878."""
        dx = self.dx
        n = (self.n_grid,) * 3
        lo = [max(0, int(np.floor(float(lo_w[i]) / dx)) - halo) for i in range(3)]
        hi = [min(n[i], int(np.ceil(float(hi_w[i]) / dx)) + halo + 1) for i in range(3)]
        dims = (hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2])
        if dims[0] <= 0 or dims[1] <= 0 or dims[2] <= 0:
            return None
        return (tuple(lo), dims)

    # ------------------------------------------------------------------ colliders
    def _register(self, kind, param, modify, aabb, label):
        self.colliders.append((kind, param))
        self.modify_bc.append(modify)
        self.collider_aabbs.append(aabb)
        self.collider_labels.append(label)

    def add_surface_collider(self, point, normal, surface="sticky", friction=0.0,
                             start_time=0.0, end_time=999.0, restitution=0.0):
        """mpm_solver_warp.py:1880-2018 (restitution > 0 also registers a rigid-body
        contact surface, :1915-1925)."""
        point = [float(v) for v in point]
        scale = 1.0 / float(f32(np.sqrt(float(sum(x ** 2 for x in normal)))))
        normal = [scale * x for x in normal]
        if restitution != 0.0:
            if not (0.0 < restitution <= 1.0):
                raise ValueError("restitution must be between 0 (exclusive) and 1 (inclusive).")
            self.rigid_surface_colliders.append({
                "point": list(point), "normal": list(normal), "friction": friction,
                "restitution": restitution, "start_time": start_time, "end_time": end_time})
        p = OPlane()
        p.start_time, p.end_time = start_time, end_time
        p.point[:] = point
        p.normal[:] = normal
        if surface == "sticky" and friction != 0:
            raise ValueError("friction must be 0 on sticky surfaces.")
        p.surface_type = {"sticky": 0, "slip": 1, "cut": 11}.get(surface, 2)
        p.friction = friction
        axis = next((i for i in range(3) if abs(abs(normal[i]) - 1.0) < 1e-6), -1)
        aabb = None
        if axis >= 0:
            p_ax, sgn = float(point[axis]), float(normal[axis])

            def aabb(axis=axis, p_ax=p_ax, sgn=sgn):
                big = 1.0e9
                lo_w = [-big] * 3
                hi_w = [big] * 3
                if sgn > 0.0:
                    hi_w[axis] = p_ax
                else:
                    lo_w[axis] = p_ax
                return self._grid_box(lo_w, hi_w, halo=1)
        self._register("plane", p, None, aabb, "plane")

    def set_velocity_on_cuboid(self, point, size, velocity, start_time=0.0, end_time=999.0,
                               reset=0):
        """mpm_solver_warp.py:2028-2106."""
        p = OBox()
        p.start_time, p.end_time = start_time, end_time
        p.point[:] = [float(v) for v in point]
        p.size[:] = [float(v) for v in size]
        p.velocity[:] = [float(v) for v in velocity]
        p.reset = reset
        force = np.zeros(3, f32)
        self._keep.append(force)
        p.force = fptr(force)
        p._force = force

        def modify(time, dt, param):
            if time >= param.start_time and time < param.end_time:
                param.point[:] = [float(f32(param.point[i] + dt * param.velocity[i]))
                                  for i in range(3)]

        def aabb(param=p):
            pt, s = param.point, param.size
            return self._grid_box((pt[0] - s[0], pt[1] - s[1], pt[2] - s[2]),
                                  (pt[0] + s[0], pt[1] + s[1], pt[2] + s[2]), halo=1)
        self._register("box", p, modify, aabb, "cuboid_velocity")

    def add_revolved_sdf_collider(self, point, rot, velocity=(0, 0, 0), omega=(0, 0, 0),
                                  outer_radius=0.089, inner_radius=0.069, half_height=0.12,
                                  inner_floor_z=-0.07, fillet_radius=0.012, friction=0.0,
                                  sticky_depth=0.0, contact_band=0.0, start_time=0.0,
                                  end_time=999.0):
        """mpm_solver_warp.py:2121-2269."""
        p = OCup()
        p.point[:] = [float(v) for v in point]
        rot_np = np.asarray(rot, dtype=np.float64).reshape(3, 3)
        p.rot[:] = [float(v) for v in rot_np.astype(f32).ravel()]
        p.velocity[:] = [float(v) for v in velocity]
        p.omega[:] = [float(v) for v in omega]
        p.start_time, p.end_time = start_time, end_time
        p.friction = friction
        p.outer_radius, p.inner_radius = outer_radius, inner_radius
        p.half_height, p.inner_floor_z, p.fillet_radius = half_height, inner_floor_z, fillet_radius
        p.sticky_depth, p.contact_band = sticky_depth, contact_band
        force, torque = np.zeros(3, f32), np.zeros(3, f32)
        self._keep += [force, torque]
        p.force, p.torque = fptr(force), fptr(torque)
        p._force, p._torque = force, torque
        idx = len(self.colliders)
        self._revolved_rot[idx] = rot_np.copy()

        def modify(time, dt, param, idx=idx):
            if time >= param.start_time and time < param.end_time:
                param.point[:] = [float(f32(param.point[i] + dt * param.velocity[i]))
                                  for i in range(3)]
                w = np.array([param.omega[0], param.omega[1], param.omega[2]], dtype=np.float64)
                w_norm = float(np.linalg.norm(w))
                ang = w_norm * dt
                if ang > 1.0e-12:
                    ax = w / w_norm
                    kx = np.array([[0.0, -ax[2], ax[1]], [ax[2], 0.0, -ax[0]],
                                   [-ax[1], ax[0], 0.0]])
                    inc = np.eye(3) + np.sin(ang) * kx + (1.0 - np.cos(ang)) * (kx @ kx)
                    self._revolved_rot[idx] = inc @ self._revolved_rot[idx]
                    param.rot[:] = [float(v) for v in self._revolved_rot[idx].astype(f32).ravel()]

        r_ball = float(np.sqrt(outer_radius ** 2 + half_height ** 2)) + float(contact_band)

        def aabb(param=p, r=r_ball):
            c = param.point
            return self._grid_box((c[0] - r, c[1] - r, c[2] - r), (c[0] + r, c[1] + r, c[2] + r),
                                  halo=1)
        self._register("cup", p, modify, aabb, "revolved_cup")
        return idx

    def set_revolved_collider_pose(self, idx, point=None, rot=None, velocity=None, omega=None):
        """mpm_solver_warp.py:2271-2285."""
        self._bc_box_cache = {}
        p = self.colliders[idx][1]
        if point is not None:
            p.point[:] = [float(v) for v in point]
        if rot is not None:
            rot_np = np.asarray(rot, dtype=np.float64).reshape(3, 3)
            self._revolved_rot[idx] = rot_np.copy()
            p.rot[:] = [float(v) for v in rot_np.astype(f32).ravel()]
        if velocity is not None:
            p.velocity[:] = [float(v) for v in velocity]
        if omega is not None:
            p.omega[:] = [float(v) for v in omega]

    def add_bounding_box(self, start_time=0.0, end_time=999.0):
        """mpm_solver_warp.py:2287-2380: six face shells sharing one param."""
        p = OWalls()
        p.start_time, p.end_time = start_time, end_time
        n, pad = self.n_grid, 3
        shells = [((0, 0, 0), (pad, n, n)), ((n - pad, 0, 0), (pad, n, n)),
                  ((0, 0, 0), (n, pad, n)), ((0, n - pad, 0), (n, pad, n)),
                  ((0, 0, 0), (n, n, pad)), ((0, 0, n - pad), (n, n, pad))]
        for box in shells:
            self._register("walls", p, None, (lambda box=box: box), "domain_walls")

    def add_sdf_collider(self, sdf_values, sdf_grads, origin, cell, center, quat=(0, 0, 0, 1),
                         velocity=(0, 0, 0), omega=(0, 0, 0), band=None, surface="separable",
                         friction=0.4, start_time=0.0, end_time=999.0):
        """mpm_solver_warp.py:2621-2777."""
        res = int(sdf_values.shape[0])
        if band is None:
            band = float(self.dx)
        sid = {"sticky": 0, "slip": 1, "separable": 2}[surface]
        vals = np.asarray(sdf_values)
        bmin = float(min(vals[0].min(), vals[-1].min(), vals[:, 0, :].min(),
                         vals[:, -1, :].min(), vals[:, :, 0].min(), vals[:, :, -1].min()))
        if band >= bmin:
            raise ValueError("SDF collider contact band reaches or exceeds the SDF grid margin")
        ext = (res - 1) * float(cell)
        corners = np.array([[float(origin[0]) + a * ext, float(origin[1]) + b * ext,
                             float(origin[2]) + c * ext]
                            for a in (0, 1) for b in (0, 1) for c in (0, 1)])
        r_max = float(np.linalg.norm(corners, axis=1).max())
        idx = len(self.colliders)
        self._sdf_guard[idx] = {"r_max": r_max, "band": float(band), "warned": False}
        p = OSdf()
        sv = np.ascontiguousarray(sdf_values, dtype=f32)
        sg = np.ascontiguousarray(sdf_grads, dtype=f32)
        self._keep += [sv, sg]
        p.sdf_val, p.sdf_grad, p.res = fptr(sv), fptr(sg), res
        p.origin[:] = [float(v) for v in origin]
        p.cell = float(cell)
        p.center[:] = [float(v) for v in center]
        p.quat[:] = [float(v) for v in quat]
        p.velocity[:] = [float(v) for v in velocity]
        p.omega[:] = [float(v) for v in omega]
        p.band, p.surface_type, p.friction = float(band), sid, float(friction)
        p.start_time, p.end_time = start_time, end_time
        force, torque = np.zeros(3, f32), np.zeros(3, f32)
        self._keep += [force, torque]
        p.force, p.torque = fptr(force), fptr(torque)
        p._force, p._torque = force, torque
        guard = self._sdf_guard[idx]

        def modify(time, dt, param):
            if time >= param.start_time and time < param.end_time:
                v = param.velocity
                w = np.array([param.omega[0], param.omega[1], param.omega[2]], dtype=float)
                speed = float(np.hypot(np.hypot(v[0], v[1]), v[2])) + \
                    float(np.linalg.norm(w)) * guard["r_max"]
                if speed * dt > guard["band"] and not guard["warned"]:
                    guard["warned"] = True
                    warnings.warn("SDF collider surface sweep exceeds the contact band: "
                                  "material can tunnel through", RuntimeWarning, stacklevel=2)
                param.center[:] = [float(f32(param.center[i] + dt * v[i])) for i in range(3)]
                q = np.array([param.quat[0], param.quat[1], param.quat[2], param.quat[3]],
                             dtype=float)
                q = _omega_step_quat(q, w, dt)
                param.quat[:] = [float(f32(t)) for t in q]

        def aabb(param=p, corners=corners):
            qx, qy, qz, qw = param.quat[0], param.quat[1], param.quat[2], param.quat[3]
            R = np.array([
                [1 - 2 * (qy * qy + qz * qz), 2 * (qx * qy - qz * qw), 2 * (qx * qz + qy * qw)],
                [2 * (qx * qy + qz * qw), 1 - 2 * (qx * qx + qz * qz), 2 * (qy * qz - qx * qw)],
                [2 * (qx * qz - qy * qw), 2 * (qy * qz + qx * qw), 1 - 2 * (qx * qx + qy * qy)]])
            c = np.array([param.center[0], param.center[1], param.center[2]])
            pts = c + corners @ R.T
            return self._grid_box(pts.min(axis=0), pts.max(axis=0), halo=1)
        self._register("sdf", p, modify, aabb, "mesh_sdf")
        return idx

    def set_sdf_pose(self, handle, center=None, quat=None, velocity=None, omega=None):
        """mpm_solver_warp.py:2779-2814 (jump warning restated)."""
        self._bc_box_cache = {}
        p = self.colliders[handle][1]
        guard = self._sdf_guard.get(handle)
        if guard is not None and not guard["warned"]:
            jump = 0.0
            if center is not None:
                jump += float(np.linalg.norm(np.array([float(center[i]) - p.center[i]
                                                       for i in range(3)])))
            if quat is not None:
                dot = abs(sum(float(quat[i]) * p.quat[i] for i in range(4)))
                jump += 2.0 * float(np.arccos(min(1.0, dot))) * guard["r_max"]
            if jump > guard["band"]:
                guard["warned"] = True
                warnings.warn("set_sdf_pose jump exceeds the contact band: tunnel risk",
                              RuntimeWarning, stacklevel=2)
        if center is not None:
            p.center[:] = [float(v) for v in center]
        if quat is not None:
            p.quat[:] = [float(v) for v in quat]
        if velocity is not None:
            p.velocity[:] = [float(v) for v in velocity]
        if omega is not None:
            p.omega[:] = [float(v) for v in omega]

    def add_point_cloud_collider(self, pts, padding=1, start_time=0.0, end_time=999.0):
        """mpm_solver_warp.py:2545-2616."""
        n = self.n_grid
        idx = np.clip((pts * self.inv_dx).astype(np.int32), 0, n - 1)
        occ = np.zeros((n, n, n), np.int32)
        occ[idx[:, 0], idx[:, 1], idx[:, 2]] = 1
        for _ in range(padding):
            dil = occ.copy()
            for ax in range(3):
                s_src = [slice(None)] * 3
                s_dst = [slice(None)] * 3
                s_src[ax], s_dst[ax] = slice(0, -1), slice(1, None)
                dil[tuple(s_dst)] |= occ[tuple(s_src)]
                s_src = [slice(None)] * 3
                s_dst = [slice(None)] * 3
                s_src[ax], s_dst[ax] = slice(1, None), slice(0, -1)
                dil[tuple(s_dst)] |= occ[tuple(s_src)]
            occ = dil
        self._keep.append(occ)
        p = OPointCloud()
        p.occupancy = iptr(occ)
        p.start_time, p.end_time = start_time, end_time
        a = np.argwhere(occ == 1)
        box = ((tuple(int(v) for v in a.min(0)),
                tuple(int(a.max(0)[i] - a.min(0)[i] + 1) for i in range(3))) if len(a) else None)
        self._register("pointcloud", p, None, (lambda box=box: box), "pointcloud")

    # ------------------------------------------------------------------ CPIC (CDF)
    def add_cdf_collider(self, cdf_values, cdf_valid, origin, cell, built_band, center,
                         quat=(0.0, 0.0, 0.0, 1.0), velocity=(0.0, 0.0, 0.0),
                         omega=(0.0, 0.0, 0.0), band=None, surface="separable", friction=0.4,
                         start_time=0.0, end_time=999.0):
        """mpm_solver_warp.py:2819-2919 (guards, AABB with a 2-node halo, lazy grids)."""
        lane = len(self.cdf_params)
        if lane >= MAX_CDF:
            raise ValueError(f"at most {MAX_CDF} CDF colliders per scene")
        res = int(cdf_values.shape[0])
        if band is None:
            band = min(float(built_band), 2.0 * float(self.dx))
        if band > float(built_band) + 1e-12:
            raise ValueError("runtime band exceeds the band the field was built with")
        if band < 1.5 * float(self.dx):
            warnings.warn("CDF band is below 1.5 dx: straddling particle-node pairs can go "
                          "untagged and the wall can leak", RuntimeWarning, stacklevel=2)
        sid = {"sticky": 0, "slip": 1, "separable": 2}[surface]
        ext = (res - 1) * float(cell)
        corners = np.array([[float(origin[0]) + a * ext, float(origin[1]) + b * ext,
                             float(origin[2]) + c * ext]
                            for a in (0, 1) for b in (0, 1) for c in (0, 1)])
        r_max = float(np.linalg.norm(corners, axis=1).max())
        self._cdf_guard[lane] = {"r_max": r_max, "band": float(band), "warned": False}
        p = OCdfCollider()
        val = np.ascontiguousarray(cdf_values, dtype=f32)
        ok = np.ascontiguousarray(cdf_valid, dtype=f32)
        self._keep += [val, ok]
        p.cdf_val, p.cdf_valid, p.res = fptr(val), fptr(ok), res
        p.origin[:] = [float(v) for v in origin]
        p.cell = float(cell)
        p.center[:] = [float(v) for v in center]
        p.quat[:] = [float(v) for v in quat]
        p.band, p.lane = float(band), lane
        p.start_time, p.end_time = start_time, end_time
        p._velocity = [float(f32(v)) for v in velocity]
        p._omega = [float(f32(v)) for v in omega]
        p._type, p._friction = sid, float(friction)
        self.cdf_params.append(p)

        def aabb(param=p, corners=corners):
            qx, qy, qz, qw = param.quat[0], param.quat[1], param.quat[2], param.quat[3]
            R = np.array([
                [1 - 2 * (qy * qy + qz * qz), 2 * (qx * qy - qz * qw), 2 * (qx * qz + qy * qw)],
                [2 * (qx * qy + qz * qw), 1 - 2 * (qx * qx + qz * qz), 2 * (qy * qz - qx * qw)],
                [2 * (qx * qz - qy * qw), 2 * (qy * qz + qx * qw), 1 - 2 * (qx * qx + qy * qy)]])
            c = np.array([param.center[0], param.center[1], param.center[2]])
            pts = c + corners @ R.T
            return self._grid_box(pts.min(axis=0), pts.max(axis=0), halo=2)
        self.cdf_aabbs.append(aabb)
        self._cdf_lane_dirty.append(True)
        self._cdf_lane_active.append(self.time >= start_time and self.time < end_time)
        self._cdf_lane_prev_box.append(None)
        if self.cdf is None:
            g = self.n_grid
            nb = (g + 3) // 4
            self.cdf_tag = np.zeros((g, g, g), np.int32)
            self.cdf_tag_prev = np.zeros((g, g, g), np.int32)
            self.cdf_d = np.full((g, g, g), 1.0e9, f32)
            self.cdf_d_prev = np.full((g, g, g), 1.0e9, f32)
            self.cdf_block = np.zeros((nb, nb, nb), np.int32)
            c = OCdf()
            c.nb[:] = [nb] * 3
            c.tag, c.tag_prev = iptr(self.cdf_tag), iptr(self.cdf_tag_prev)
            c.d, c.d_prev = fptr(self.cdf_d), fptr(self.cdf_d_prev)
            c.block = iptr(self.cdf_block)
            self.cdf = c
        self.cdf.n_cdf = len(self.cdf_params)
        self._sync_cdf_lane(lane)
        return lane

    def _sync_cdf_lane(self, lane):
        """mpm_solver_warp.py:2921-2943: pose / contact state into the current lane arrays."""
        p, c = self.cdf_params[lane], self.cdf
        c.lane_center[lane][:] = [p.center[0], p.center[1], p.center[2]]
        c.lane_velocity[lane][:] = p._velocity
        c.lane_omega[lane][:] = p._omega
        c.lane_friction[lane] = p._friction
        c.lane_type[lane] = p._type

    def set_cdf_pose(self, handle, center=None, quat=None, velocity=None, omega=None):
        """mpm_solver_warp.py:2945-2996."""
        p = self.cdf_params[handle]
        guard = self._cdf_guard.get(handle)
        if guard is not None and not guard["warned"]:
            jump = 0.0
            if center is not None:
                jump += float(np.linalg.norm(np.array([float(center[i]) - p.center[i]
                                                       for i in range(3)])))
            if quat is not None:
                dot = abs(sum(float(quat[i]) * p.quat[i] for i in range(4)))
                jump += 2.0 * float(np.arccos(min(1.0, dot))) * guard["r_max"]
            if jump > guard["band"]:
                guard["warned"] = True
                warnings.warn("set_cdf_pose jumped the surface by more than the band in one "
                              "tick", RuntimeWarning, stacklevel=2)
        moved = synced = False
        if center is not None:
            new = [float(f32(v)) for v in center]
            moved |= new != [p.center[0], p.center[1], p.center[2]]
            p.center[:] = new
        if quat is not None:
            new = [float(f32(v)) for v in quat]
            moved |= new != [p.quat[i] for i in range(4)]
            p.quat[:] = new
        if velocity is not None:
            new = [float(f32(v)) for v in velocity]
            synced |= new != p._velocity
            p._velocity = new
        if omega is not None:
            new = [float(f32(v)) for v in omega]
            synced |= new != p._omega
            p._omega = new
        if moved:
            self._cdf_lane_dirty[handle] = True
        if moved or synced:
            self._sync_cdf_lane(handle)
            self._cdf_lane_prev_stale = True

    def _cdf_dirty_box(self):
        """mpm_solver_warp.py:2998-3018."""
        boxes = []
        for lane, fn in enumerate(self.cdf_aabbs):
            if not self._cdf_lane_dirty[lane]:
                continue
            b = fn()
            prev = self._cdf_lane_prev_box[lane]
            if b is not None:
                boxes.append(b)
            if prev is not None:
                boxes.append(prev)
            self._cdf_lane_prev_box[lane] = b
        if not boxes:
            return None
        lo = tuple(min(b[0][i] for b in boxes) for i in range(3))
        hi = tuple(max(b[0][i] + b[1][i] for b in boxes) for i in range(3))
        return (lo, tuple(hi[i] - lo[i] for i in range(3)))

    @staticmethod
    def _cdf_box_intersect(a, b):
        """mpm_solver_warp.py:3020-3028."""
        if a is None or b is None:
            return None
        lo = tuple(max(a[0][i], b[0][i]) for i in range(3))
        hi = tuple(min(a[0][i] + a[1][i], b[0][i] + b[1][i]) for i in range(3))
        if any(hi[i] <= lo[i] for i in range(3)):
            return None
        return (lo, tuple(hi[i] - lo[i] for i in range(3)))

    def _cdf_reconcile_prev(self):
        """mpm_solver_warp.py:3030-3038."""
        if self._cdf_tag_prev_stale_box is not None:
            lo, dims = self._cdf_tag_prev_stale_box
            lib().o_cdf_copy_prev(self._state(), self._model(), i3(lo), i3(dims))
            self._cdf_tag_prev_stale_box = None

    def _stamp_cdf(self, dt, lag_prev=False):
        """mpm_solver_warp.py:3040-3098: per-substep restamp of the dirty lanes' boxes."""
        if self.cdf is None or self.cdf.n_cdf == 0:
            return
        self._cdf_reconcile_prev()
        for lane, p in enumerate(self.cdf_params):
            active = self.time >= p.start_time and self.time < p.end_time
            if active != self._cdf_lane_active[lane]:
                self._cdf_lane_active[lane] = active
                self._cdf_lane_dirty[lane] = True
        if self._cdf_force_stamp:
            self._cdf_lane_dirty = [True] * len(self.cdf_params)
        if not any(self._cdf_lane_dirty):
            return
        box = self._cdf_dirty_box()
        self._cdf_lane_dirty = [False] * len(self.cdf_params)
        if box is None:
            return
        self._cdf_stamp_count += 1
        L = lib()
        st, md = self._state(), self._model()
        lo, dims = box
        L.o_cdf_zero_tags(st, md, i3(lo), i3(dims))
        for lane, p in enumerate(self.cdf_params):
            clip = self._cdf_box_intersect(box, self.cdf_aabbs[lane]())
            if clip is None:
                continue
            L.o_cdf_stamp(float(f32(self.time)), st, md, C.byref(p), i3(clip[0]), i3(clip[1]))
        if lag_prev:
            self._cdf_tag_prev_stale_box = box
        else:
            L.o_cdf_copy_prev(st, md, i3(lo), i3(dims))
        blo = tuple(lo[i] >> 2 for i in range(3))
        bdims = tuple(((lo[i] + dims[i] - 1) >> 2) - blo[i] + 1 for i in range(3))
        L.o_cdf_block_mask(st, md, i3(blo), i3(bdims))

    def _cdf_copy_lane_prev(self):
        """mpm_solver_warp.py:3100-3105."""
        c = self.cdf
        for lane in range(MAX_CDF):
            c.lane_center_prev[lane][:] = list(c.lane_center[lane])
            c.lane_velocity_prev[lane][:] = list(c.lane_velocity[lane])
            c.lane_omega_prev[lane][:] = list(c.lane_omega[lane])
        self._cdf_lane_prev_stale = False

    def _integrate_cdf_poses(self, dt, lag_prev=False):
        """mpm_solver_warp.py:3107-3148: advance moving lanes one substep."""
        if not self.cdf_params:
            return
        if lag_prev and self._cdf_lane_prev_stale:
            self._cdf_copy_lane_prev()
        for lane, p in enumerate(self.cdf_params):
            if not (self.time >= p.start_time and self.time < p.end_time):
                continue
            v = p._velocity
            w = np.array(p._omega, dtype=float)
            if v[0] == 0.0 and v[1] == 0.0 and v[2] == 0.0 and not w.any():
                continue
            guard = self._cdf_guard[lane]
            speed = float(np.hypot(np.hypot(v[0], v[1]), v[2])) + \
                float(np.linalg.norm(w)) * guard["r_max"]
            if speed * dt > guard["band"] and not guard["warned"]:
                guard["warned"] = True
                warnings.warn("CDF collider surface sweeps more than the band per substep: "
                              "material can end up on the wrong side", RuntimeWarning,
                              stacklevel=2)
            c = [p.center[0], p.center[1], p.center[2]]
            p.center[:] = [c[i] + dt * v[i] for i in range(3)]
            q = np.array([p.quat[0], p.quat[1], p.quat[2], p.quat[3]], dtype=float)
            q = _omega_step_quat(q, w, dt)
            p.quat[:] = [float(t) for t in q]
            self._sync_cdf_lane(lane)
            self._cdf_lane_dirty[lane] = True
            self._cdf_lane_prev_stale = True
        if not lag_prev and self._cdf_lane_prev_stale:
            self._cdf_copy_lane_prev()

    def reset_cdf_wrench(self):
        """mpm_solver_warp.py:3153-3155."""
        if self.cdf is not None:
            for lane in range(MAX_CDF):
                self.cdf.reaction_force[lane][:] = [0.0, 0.0, 0.0]
                self.cdf.reaction_torque[lane][:] = [0.0, 0.0, 0.0]

    # ------------------------------------------------------------------ BC application
    _COLLIDE = {"plane": "o_collide_plane", "box": "o_collide_box", "cup": "o_collide_cup",
                "walls": "o_collide_walls", "sdf": "o_collide_sdf",
                "pointcloud": "o_collide_pointcloud"}

    def _apply_grid_bc(self, dt):
        """mpm_solver_warp.py:1283-1336."""
        full = ((0, 0, 0), (self.n_grid,) * 3)
        st, md = self._state(), self._model()
        for k, (kind, param) in enumerate(self.colliders):
            lo, dims = full
            skip = False
            fn = self.collider_aabbs[k]
            if self.restrict_bc and fn is not None:
                static = self.modify_bc[k] is None
                if static and k in self._bc_box_cache:
                    skip, lo, dims = self._bc_box_cache[k]
                else:
                    box = fn()
                    if box is None:
                        skip = True
                    else:
                        lo, dims = box
                    if static:
                        self._bc_box_cache[k] = (skip, lo, dims)
            if not skip:
                getattr(lib(), self._COLLIDE[kind])(float(f32(self.time)), dt, st, md,
                                                    C.byref(param), i3(lo), i3(dims))
            if self.modify_bc[k] is not None:
                self.modify_bc[k](self.time, dt, param)
        self._integrate_cdf_poses(dt, lag_prev=self._cdf_lag_prev)

    # ------------------------------------------------------------------ substeps
    def _boxes(self):
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

    def p2g2p(self, dt):
        """Split substep, mpm_solver_warp.py:1092-1281 (CPU path: no graphs)."""
        L = lib()
        self._cdf_lag_prev = False
        self._stamp_cdf(dt)
        st, md = self._state(), self._model()
        sparse = self.sparse_ready and self.restrict_grid
        if sparse:
            self._prev_grid_box = None
            b = self._blk
            if b["union_n"] > 0:
                L.o_blocks_grid_op(st, md, 0, dt, 0.0, iptr(b["union_list"]), b["union_n"],
                                   i3(b["bd"]))
        else:
            g, z = self._boxes()
            L.o_zero_grid(st, md, i3(z[0]), i3(z[1]), 7)
        self._apply_particle_modifiers(dt)
        L.o_compute_stress(st, md, dt)
        L.o_p2g(st, md, dt)
        if sparse:
            b = self._blk
            if b["cur_n"] > 0:
                L.o_blocks_grid_op(st, md, 1, dt, 0.0, iptr(b["cur_list"]), b["cur_n"],
                                   i3(b["bd"]))
        else:
            L.o_grid_normalize(st, md, dt, i3(g[0]), i3(g[1]))
        if self.grid_v_damping_scale < 1.0:
            if sparse:
                b = self._blk
                if b["cur_n"] > 0:
                    L.o_blocks_grid_op(st, md, 2, dt, self.grid_v_damping_scale,
                                       iptr(b["cur_list"]), b["cur_n"], i3(b["bd"]))
            else:
                L.o_grid_damping(st, md, self.grid_v_damping_scale, i3(g[0]), i3(g[1]))
        self._apply_grid_bc(dt)
        L.o_g2p(self._state(), self._model(), dt)
        self._p2g2p_tail(dt)

    def p2g2p_fused_tick(self, dt, substeps):
        """Fused tick, mpm_solver_warp.py:1379-1546 (CPU path: no graphs)."""
        L = lib()
        for s in range(substeps):
            self._cdf_lag_prev = True
            self._stamp_cdf(dt, lag_prev=True)
            st, md = self._state(), self._model()
            g, z = self._boxes()
            if s == 0:
                L.o_zero_grid(st, md, i3(z[0]), i3(z[1]), 7)
                L.o_compute_stress(st, md, dt)
                L.o_p2g(st, md, dt)
            else:
                L.o_zero_grid(st, md, i3(z[0]), i3(z[1]), 3)
                L.o_g2p_stress_p2g(st, md, dt)
                L.o_zero_grid(st, md, i3(z[0]), i3(z[1]), 4)
            L.o_grid_normalize(st, md, dt, i3(g[0]), i3(g[1]))
            if self.grid_v_damping_scale < 1.0:
                L.o_grid_damping(st, md, self.grid_v_damping_scale, i3(g[0]), i3(g[1]))
            self._apply_grid_bc(dt)
            self.time = self.time + dt
        if self.cdf is not None and self.cdf.n_cdf > 0:
            self._cdf_reconcile_prev()
        L.o_g2p(self._state(), self._model(), dt)

    # ------------------------------------------------------------------ sparse blocks
    def rebuild_active_blocks(self):
        """mpm_solver_warp.py:989-1026."""
        if self._blk is None:
            n = self.n_grid
            bd = (int(np.ceil(n / 4)),) * 3
            nb = bd[0] * bd[1] * bd[2]
            self._blk = {"bd": bd, "raw": np.zeros(bd, np.int32), "cur": np.zeros(bd, np.int32),
                         "prev": np.zeros(bd, np.int32), "cur_list": np.zeros(nb, np.int32),
                         "union_list": np.zeros(nb, np.int32), "cur_n": 0, "union_n": 0}
        b = self._blk
        b["prev"], b["cur"] = b["cur"], b["prev"]
        b["raw"][:] = 0
        L = lib()
        L.o_blocks_mark(fptr(self.x), self.n_particles, self.inv_dx, i3(b["bd"]), iptr(b["raw"]))
        L.o_blocks_dilate(iptr(b["raw"]), i3(b["bd"]), iptr(b["cur"]))
        counts = np.zeros(2, np.int32)
        L.o_blocks_compact(iptr(b["cur"]), iptr(b["prev"]), i3(b["bd"]), iptr(b["cur_list"]),
                           iptr(b["union_list"]), iptr(counts))
        b["cur_n"], b["union_n"] = int(counts[0]), int(counts[1])
        self.sparse_ready = True

    def active_block_fraction(self):
        if not self.sparse_ready:
            return 1.0
        bd = self._blk["bd"]
        return self._blk["cur_n"] / float(bd[0] * bd[1] * bd[2])

    def permute_particles(self, perm):
        """mpm_solver_warp.py:1548-1619 (every per-particle array, incl. mu/lam/yield)."""
        perm = np.asarray(perm)
        for nm in ("x", "v", "F", "F_trial", "R", "stress", "C", "L", "vol", "mass",
                   "density", "Jp", "mu", "lam", "yield_stress", "selection", "material",
                   "particle_cdf_tag", "rigid_id", "x_ref"):
            a = getattr(self, nm)
            a[...] = a[perm]
        for nm in ("init_cov", "cov"):
            a = getattr(self, nm).reshape(-1, 6)
            a[...] = a[perm]
        for p in self.impulse_params + self.particle_velocity_modifier_params:
            p["mask"] = p["mask"][perm]

    def compute_R(self):
        lib().o_compute_R(self._state())
        return self.R

    def compute_cov(self):
        lib().o_compute_cov(self._state())
        return self.cov


class OracleSolver:
    """Restatement of the typed wrapper Solver (src/warpmpm/core/solver.py:57-642) over
    OracleSim. Accepts the same Material objects (anything with .resolve())."""

    def __init__(self, grid, guard_interval=1, sparse=False, fused=True, sort_interval=0,
                 periodic_x=False):
        self.grid = grid
        self.guard_interval = guard_interval
        self.sparse = sparse
        self.fused = fused
        self.sort_interval = sort_interval
        self.periodic_x = periodic_x
        self._sim = None
        self._step = 0
        self._tick = 0

    def load_particles(self, pos, vol, cov=None, cov_mode="step"):
        self._vol0 = np.asarray(vol, f32).copy()
        self._sim = OracleSim(pos, vol, self.grid.n_grid, self.grid.grid_lim, cov=cov)
        if self.periodic_x:
            self._sim.periodic_x = 1
        if cov is not None and cov_mode == "step":
            self._sim.update_cov_with_F = 1
            self._sim.cov[:] = self._sim.init_cov
        return self

    def set_material(self, material, **overrides):
        if hasattr(material, "resolve"):
            name, params = material.resolve()
        else:
            name, params = str(material), {}
        params = {**params, **overrides}
        self._sim.set_parameters_dict({"material": name, "g": [0.0, 0.0, -9.81], **params})
        self._sim.finalize_mu_lam()
        return self

    def set_material_range(self, start, end, material, **overrides):
        if hasattr(material, "resolve"):
            name, params = material.resolve()
        else:
            name, params = str(material), {}
        self._sim.set_parameters_for_particles(start, end,
                                               {"material": name, **params, **overrides})
        return self

    def add_plane(self, point, normal, surface="sticky", friction=0.0, restitution=0.0):
        self._sim.add_surface_collider(tuple(point), tuple(normal), surface, friction=friction,
                                       restitution=restitution)
        return self

    def finalize_rigid_bodies(self):
        self._sim.finalize_rigid_bodies()
        return self

    def rigid_state(self, body=0):
        s = self._sim
        return {"com": s.rigid_x_cm[body].copy(), "v": s.rigid_v_cm[body].copy(),
                "omega": s.rigid_omega[body].copy(), "R": s.rigid_orientation[body].copy()}

    def add_box(self, center, half_size, velocity=(0.0, 0.0, 0.0), start_time=0.0,
                end_time=1.0e9):
        self._sim.set_velocity_on_cuboid(tuple(center), tuple(half_size), tuple(velocity),
                                         start_time=start_time, end_time=end_time)
        return len(self._sim.colliders) - 1

    def set_box(self, handle, center=None, velocity=None):
        p = self._sim.colliders[handle][1]
        self._sim._bc_box_cache = {}
        if center is not None:
            p.point[:] = [float(v) for v in center]
        if velocity is not None:
            p.velocity[:] = [float(v) for v in velocity]
        return self

    def reset_tool_force(self, handle):
        self._sim.colliders[handle][1]._force[:] = 0
        return self

    def tool_force(self, handle, dt):
        return np.asarray(self._sim.colliders[handle][1]._force, dtype=float) / dt

    def add_cup(self, profile, center, quat=(1.0, 0.0, 0.0, 0.0), velocity=(0, 0, 0),
                omega=(0, 0, 0), friction=0.05, sticky_cells=1.5, contact_cells=0.5,
                start_time=0.0, end_time=1.0e9):
        sticky = min(sticky_cells * self.grid.dx, 0.45 * profile.wall_thickness)
        return self._sim.add_revolved_sdf_collider(
            tuple(center), quat_to_mat(quat), velocity=tuple(velocity), omega=tuple(omega),
            outer_radius=profile.outer_radius, inner_radius=profile.inner_radius,
            half_height=profile.half_height, inner_floor_z=profile.inner_floor_z,
            fillet_radius=profile.fillet_radius, friction=friction, sticky_depth=sticky,
            contact_band=contact_cells * self.grid.dx, start_time=start_time,
            end_time=end_time)

    def set_cup(self, handle, center=None, quat=None, velocity=None, omega=None):
        rot = None if quat is None else quat_to_mat(quat)
        self._sim.set_revolved_collider_pose(handle, point=center, rot=rot,
                                             velocity=velocity, omega=omega)
        return self

    def reset_cup_wrench(self, handle):
        p = self._sim.colliders[handle][1]
        p._force[:] = 0
        p._torque[:] = 0
        return self

    def cup_wrench(self, handle, dt):
        p = self._sim.colliders[handle][1]
        return {"force": np.asarray(p._force, float) / dt, "torque": np.asarray(p._torque, float) / dt}

    reset_sdf_force = reset_cup_wrench
    sdf_wrench = cup_wrench

    def add_domain_walls(self, start_time=0.0, end_time=1.0e9):
        self._sim.add_bounding_box(start_time=start_time, end_time=end_time)
        return self

    def add_sdf_collider(self, sdf, center, quat=(0.0, 0.0, 0.0, 1.0), velocity=(0, 0, 0),
                         omega=(0, 0, 0), band=None, surface="separable", friction=0.4,
                         start_time=0.0, end_time=1.0e9):
        return self._sim.add_sdf_collider(sdf.values, sdf.grads, sdf.origin, sdf.cell, center,
                                          quat=quat, velocity=velocity, omega=omega, band=band,
                                          surface=surface, friction=friction,
                                          start_time=start_time, end_time=end_time)

    def set_sdf_pose(self, handle, center=None, quat=None, velocity=None, omega=None):
        self._sim.set_sdf_pose(handle, center=center, quat=quat, velocity=velocity, omega=omega)
        return self

    def add_cdf_collider(self, cdf, center, quat=(0.0, 0.0, 0.0, 1.0), velocity=(0.0, 0.0, 0.0),
                         omega=(0.0, 0.0, 0.0), band=None, surface="separable", friction=0.4,
                         start_time=0.0, end_time=1.0e9):
        """solver.py:363-386."""
        if self.periodic_x:
            raise NotImplementedError("CDF colliders do not wrap the periodic x axis")
        return self._sim.add_cdf_collider(cdf.values, cdf.valid, cdf.origin, cdf.cell, cdf.band,
                                          center, quat=quat, velocity=velocity, omega=omega,
                                          band=band, surface=surface, friction=friction,
                                          start_time=start_time, end_time=end_time)

    def set_cdf_pose(self, handle, center=None, quat=None, velocity=None, omega=None):
        self._sim.set_cdf_pose(handle, center=center, quat=quat, velocity=velocity, omega=omega)
        return self

    def reset_cdf_wrench(self):
        self._sim.reset_cdf_wrench()
        return self

    def cdf_wrench(self, handle, dt):
        """solver.py:401-413: blocked-deposit reaction since the last reset, / dt."""
        c = self._sim.cdf
        f = np.array(list(c.reaction_force[handle]), dtype=float)
        t = np.array(list(c.reaction_torque[handle]), dtype=float)
        return {"force": f / dt, "torque": t / dt}

    def step(self, dt, substeps=1):
        """solver.py:429-452."""
        if self._tick % max(1, self.guard_interval) == 0:
            self._update_grid_box(dt, substeps)
        if self.sort_interval and self._tick % self.sort_interval == 0:
            self._sort_particles()
        if self.sparse:
            self._sim.rebuild_active_blocks()
        self._tick += 1
        sim = self._sim
        if (self.fused and not self.sparse and sim.n_rigid_bodies == 0 and not sim.impulse_params
                and not sim.particle_velocity_modifier_params):
            self._sim.p2g2p_fused_tick(dt, substeps)
            self._step += substeps
        else:
            for _ in range(substeps):
                self._sim.p2g2p(dt)
                self._step += 1
        return self

    def _sort_particles(self):
        """solver.py:454-467."""
        x = self.x()
        base = (np.floor(x / self.grid.dx - 0.5).astype(np.int64)) >> 2
        nb = (self.grid.n_grid >> 2) + 2
        keys = (base[:, 0] * nb + base[:, 1]) * nb + base[:, 2]
        if np.all(np.diff(keys) >= 0):
            return False
        perm = np.argsort(keys, kind="stable")
        self._sim.permute_particles(perm)
        self._vol0 = self._vol0[perm]
        return True

    def _update_grid_box(self, dt, substeps):
        """solver.py:485-518."""
        x, v = self.x(), self.v()
        dx = self.grid.dx
        lim = self.grid.grid_lim
        g = x[:, 1:] if self.periodic_x else x
        if g.min() < 1.5 * dx or g.max() > lim - 2.5 * dx:
            raise RuntimeError("particles within 2 cells of the grid edge: the P2G stencil "
                               "would write out of bounds")
        pad = 3.0 * dx + 1.5 * float(np.abs(v).max()) * dt * substeps
        lo = x.min(0) - pad
        hi = x.max(0) + pad
        if self.periodic_x:
            lo[0], hi[0] = 0.0, lim
        self._sim.grid_launch_box = self._sim._grid_box(lo, hi, halo=0)

    def set_x(self, pos):
        """solver.py:520-526"""
        self._sim.x[:] = np.asarray(pos, np.float32)
        return self

    def set_v(self, vel):
        """solver.py:528-534"""
        self._sim.v[:] = np.asarray(vel, np.float32)
        return self

    def x(self):
        return self._sim.x.copy()

    def v(self):
        return self._sim.v.copy()

    def F(self):
        return self._sim.F.copy()

    def stress(self):
        return self._sim.stress.copy()

    def L(self):
        return self._sim.L.copy()

    @property
    def n_particles(self):
        return 0 if self._sim is None else self._sim.n_particles
