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
