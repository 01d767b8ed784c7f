"""Quadtree node splitting (``_discretization_tree_operations_2D.py`` of the
reference): ``add_four_children`` and ``get_four_children``."""

from ._adaptive import add_four_children, get_four_children

__all__ = ["add_four_children", "get_four_children"]
