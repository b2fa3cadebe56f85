"""Collider geometry helpers (host side)."""
from .glass import (GlassProfile, angular_velocity_between, cavity_mask, cup_fill, glass_cavity_profile,
                    glass_cdf, glass_sdf,
                    glass_sdf_local, quat_conj, quat_from_axis_angle, quat_mul, quat_to_mat,
                    solid_mask, world_to_local)

__all__ = ["GlassProfile", "angular_velocity_between", "cavity_mask", "cup_fill", "glass_cavity_profile",
           "glass_cdf", "glass_sdf",
           "glass_sdf_local", "quat_conj", "quat_from_axis_angle", "quat_mul", "quat_to_mat",
           "solid_mask", "world_to_local"]
