"""Octree node splitting (``_discretization_tree_operations_3D.py`` of the
reference): ``add_eight_children`` and ``get_eight_children``."""

from ._adaptive3d import add_eight_children, get_eight_children

__all__ = ["add_eight_children", "get_eight_children"]
