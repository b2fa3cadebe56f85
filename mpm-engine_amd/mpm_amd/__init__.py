"""mpm_amd: warpmpm's explicit MLS-MPM/APIC substep on AMD MI355X (gfx950).

Drop-in for ``warpmpm``'s typed API (Solver, GridConfig, materials, colliders, scenes)
with ``device="hip"``; the substep runs in hand-written HIP kernels (libmpm_hip.so)
behind the C ABI in include/mpm_hip.h.
"""
from .colliders.glass import GlassProfile, cup_fill
from .core.solver import GridConfig, Solver
from .materials import Material, elastic, granular, newtonian, tabulated_viscous, vonmises
from .scenes import block, dough

__all__ = ["GlassProfile", "GridConfig", "Material", "Solver", "block", "cup_fill", "dough",
           "elastic", "granular", "newtonian", "tabulated_viscous", "vonmises"]
