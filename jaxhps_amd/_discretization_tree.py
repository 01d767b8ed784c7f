"""Quadtree / octree node types.

Mirrors ``_discretization_tree.py:77-161`` (2D) and ``:164-264`` (3D) of the
reference: plain host objects holding the box bounds, depth and children.
Only uniform trees are used by this package, so the per-node solver data
fields of the reference (``NodeData``) are kept only as an empty container
for interface compatibility.
"""

from __future__ import annotations

from typing import Tuple


class NodeData:
    """Per-node container (``_discretization_tree.py:6-33``)."""

    def __init__(self):
        self.T = None
        self.h = None
        self.S = None
        self.g_tilde = None
        self.Y = None
        self.v = None
        self.u = None
        self.g = None
        self.L_4f1 = None
        self.L_1f4 = None
        self.l2_nrm = 0.0


class DiscretizationNode2D:
    """Axis-aligned rectangle [xmin, xmax] x [ymin, ymax]."""

    def __init__(
        self,
        xmin: float,
        xmax: float,
        ymin: float,
        ymax: float,
        depth: int = 0,
        children: Tuple["DiscretizationNode2D", ...] = (),
    ):
        self.xmin = xmin
        self.xmax = xmax
        self.ymin = ymin
        self.ymax = ymax
        self.depth = depth
        self.data = NodeData()
        self.n_0 = None
        self.n_1 = None
        self.n_2 = None
        self.n_3 = None
        self.children = children

    def __repr__(self) -> str:
        return (
            f"DiscretizationNode2D(xmin={self.xmin}, xmax={self.xmax}, "
            f"ymin={self.ymin}, ymax={self.ymax}, depth={self.depth})"
        )


class DiscretizationNode3D:
    """Axis-aligned box [xmin, xmax] x [ymin, ymax] x [zmin, zmax]."""

    def __init__(
        self,
        xmin: float,
        xmax: float,
        ymin: float,
        ymax: float,
        zmin: float,
        zmax: float,
        depth: int = 0,
        children: Tuple["DiscretizationNode3D", ...] = (),
    ):
        self.xmin = xmin
        self.xmax = xmax
        self.ymin = ymin
        self.ymax = ymax
        self.zmin = zmin
        self.zmax = zmax
        self.depth = depth
        self.data = NodeData()
        self.n_0 = None
        self.n_1 = None
        self.n_2 = None
        self.n_3 = None
        self.n_4 = None
        self.n_5 = None
        self.children = children

    def __repr__(self) -> str:
        return (
            f"DiscretizationNode3D(xmin={self.xmin}, xmax={self.xmax}, "
            f"ymin={self.ymin}, ymax={self.ymax}, zmin={self.zmin}, "
            f"zmax={self.zmax}, depth={self.depth})"
        )


def get_all_leaves(node):
    """All leaf descendants in child order (``_discretization_tree.py:282-302``)."""
    if not len(node.children):
        return (node,)
    out = ()
    for child in node.children:
        out += get_all_leaves(child)
    return out


def get_discretization_node_area(node) -> float:
    if isinstance(node, DiscretizationNode3D):
        return (node.xmax - node.xmin) * (node.ymax - node.ymin) * (node.zmax - node.zmin)
    return (node.xmax - node.xmin) * (node.ymax - node.ymin)


def get_depth(node) -> int:
    return max(leaf.depth for leaf in get_all_leaves(node))


def get_nodes_at_level(node, level: int):
    if node.depth == level:
        return (node,)
    out = ()
    for child in node.children:
        out += get_nodes_at_level(child, level)
    return out
