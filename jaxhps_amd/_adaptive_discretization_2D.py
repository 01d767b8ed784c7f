"""Adaptive refinement of 2D trees (``_adaptive_discretization_2D.py`` of the
reference)."""

from ._adaptive import (
    find_or_add_child_2D,
    generate_adaptive_mesh_level_restriction_2D,
    get_squared_l2_norm_single_panel,
    node_to_bounds,
    patches_to_check,
)

__all__ = [
    "generate_adaptive_mesh_level_restriction_2D",
    "get_squared_l2_norm_single_panel",
    "node_to_bounds",
    "patches_to_check",
    "find_or_add_child_2D",
]
