"""Adaptive refinement of 3D trees (``_adaptive_discretization_3D.py`` of the
reference)."""

from ._adaptive3d import (
    find_or_add_child,
    generate_adaptive_mesh_level_restriction,
    get_squared_l2_norm_single_voxel,
    volumes_to_check,
)

__all__ = [
    "generate_adaptive_mesh_level_restriction",
    "get_squared_l2_norm_single_voxel",
    "volumes_to_check",
    "find_or_add_child",
]
