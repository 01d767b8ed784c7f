"""Leaf adjacency on the Gauss boundary nodes of a uniform quadtree.

Leaf l of a level-L uniform tree sits at grid cell (ix, iy) given by its
base-4 child digits (SW=0, SE=1, NE=2, NW=3, most significant first; the
order of ``_grid_creation_2D.py:35-46`` subdivision).  Its 4q Gauss boundary
slots run S (W->E), E (S->N), N (E->W), W (N->S), q per side
(``_grid_creation_2D.py:77-124`` for the root boundary, the same walk per
leaf).  Two leaves sharing an edge see its Gauss nodes in opposite orders.
"""

from __future__ import annotations

import numpy as np


def leaf_cells_2D(L: int):
    """(ix, iy) of every leaf in tree order."""
    nl = 4**L
    idx = np.arange(nl)
    ix = np.zeros(nl, dtype=np.int64)
    iy = np.zeros(nl, dtype=np.int64)
    for lvl in range(L):
        d = (idx // 4 ** (L - 1 - lvl)) % 4
        ix = 2 * ix + ((d == 1) | (d == 2))
        iy = 2 * iy + ((d == 2) | (d == 3))
    return ix, iy


def gauss_neighbors_2D(L: int, q: int) -> np.ndarray:
    """(n_leaves, 4q) int32: neighbour slot l'*4q + j' across the shared edge,
    or -(root boundary index + 1) on the root boundary (root order S, E, N, W
    as Domain.boundary_points)."""
    n = 2**L
    nq = 4 * q
    ix, iy = leaf_cells_2D(L)
    cell_to_leaf = np.empty((n, n), dtype=np.int64)
    cell_to_leaf[ix, iy] = np.arange(4**L)
    k = np.arange(q)
    rk = q - 1 - k
    nside = n * q
    out = np.empty((4**L, nq), dtype=np.int64)
    for l in range(4**L):
        x, y = ix[l], iy[l]
        # S side; neighbour below sees the edge as its N side
        out[l, 0:q] = -(x * q + k + 1) if y == 0 else cell_to_leaf[x, y - 1] * nq + 2 * q + rk
        # E side; neighbour to the right sees it as its W side
        out[l, q:2 * q] = -(nside + y * q + k + 1) if x == n - 1 else cell_to_leaf[x + 1, y] * nq + 3 * q + rk
        # N side (E->W); root N runs E->W as well
        out[l, 2 * q:3 * q] = (-(2 * nside + (n - 1 - x) * q + k + 1) if y == n - 1
                               else cell_to_leaf[x, y + 1] * nq + 0 * q + rk)
        # W side (N->S)
        out[l, 3 * q:4 * q] = (-(3 * nside + (n - 1 - y) * q + k + 1) if x == 0
                               else cell_to_leaf[x - 1, y] * nq + 1 * q + rk)
    return out.astype(np.int32)


def gauss_neighbors_3D(root, L: int, q: int) -> np.ndarray:
    """(n_leaves, 6q^2) int32 for a uniform octree: the neighbour slot
    l'*6q^2 + j' holding the same Gauss node across a shared face, or
    -(root boundary index + 1) on the root boundary.  A leaf's slots are its
    faces 1..6 (x-, x+, y-, y+, z-, z+) with a q x q Gauss grid each, the
    order of compute_boundary_Gauss_points_uniform_3D on the leaf's box; the
    root order is Domain.boundary_points.  Nodes are matched by coordinates."""
    from ._discretization_tree import DiscretizationNode3D
    from ._grid_creation import _subdivide_3D, compute_boundary_Gauss_points_uniform_3D

    boxes = _subdivide_3D([root.xmin, root.xmax, root.ymin, root.ymax, root.zmin, root.zmax], L)
    nl = boxes.shape[0]
    nq = 6 * q * q
    leaf_pts = np.concatenate(
        [compute_boundary_Gauss_points_uniform_3D(DiscretizationNode3D(*b), 0, q) for b in boxes])
    root_pts = compute_boundary_Gauss_points_uniform_3D(root, L, q)
    scale = max(root.xmax - root.xmin, root.ymax - root.ymin, root.zmax - root.zmin)
    keys = np.round(np.concatenate([leaf_pts, root_pts]) / scale * 2**40).astype(np.int64)
    _, group = np.unique(keys, axis=0, return_inverse=True)
    group = group.ravel()
    ids = np.concatenate([np.arange(nl * nq), -(np.arange(root_pts.shape[0]) + 1)])
    order = np.argsort(group, kind="stable")
    g_sorted = group[order]
    if np.any(np.bincount(g_sorted) != 2):
        raise ValueError("Gauss nodes of a uniform octree must pair up (two leaves, or a leaf and the root boundary)")
    a, b = ids[order[0::2]], ids[order[1::2]]
    out = np.empty(nl * nq, dtype=np.int64)
    # every pair holds at least one leaf slot (a root node is matched by exactly one leaf)
    out[a[a >= 0]] = b[a >= 0]
    out[b[b >= 0]] = a[b >= 0]
    return out.reshape(nl, nq).astype(np.int32)
