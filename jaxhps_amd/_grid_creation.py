"""Uniform leaf grids, boundary Gauss points and node orderings (host, float64).

The orderings here are the contract every kernel relies on; they follow the
reference exactly:

* 2D leaf point order: ``_grid_creation_2D.py:53-68`` (meshgrid(x, flipud(y),
  indexing="ij")) permuted by ``rearrange_indices_ext_int`` (:204-233) so the
  4(p-1) boundary points come first, counter-clockwise from the SW corner.
* 2D sibling order a=SW, b=SE, c=NE, d=NW (:35-46), applied level by level.
* 2D root boundary Gauss order S(W->E), E(S->N), N(E->W), W(N->S) (:77-124).
* 3D leaf order (``_grid_creation_3D.py:80-116``), octant order a-d upper z,
  e-h lower z (:29-75), faces x-, x+, y-, y+, z-, z+ (:282-319, :119-190).
"""

from __future__ import annotations

import numpy as np

from .quadrature import affine_transform, chebyshev_points, gauss_points


# --------------------------------------------------------------------------
# 2D
# --------------------------------------------------------------------------
def rearrange_indices_ext_int_2D(n: int) -> np.ndarray:
    """Permutation of the n x n ij-flattened Chebyshev grid putting the
    boundary first (``_grid_creation_2D.py:204-233``).

    Flat index k = i*n + j addresses x_i (ascending) and y_{n-1-j} (j = 0 is
    the north edge).  Boundary sequence: S edge W->E (n points), E edge S->N
    without SE (n-1), N edge E->W without NE ending at NW (n-1), W edge N->S
    without either corner (n-2); then the interior in ascending flat index.
    """
    south = [i * n + (n - 1) for i in range(n)]
    east = [(n - 1) * n + j for j in range(n - 2, -1, -1)]
    north = [i * n for i in range(n - 2, 0, -1)] + [0]
    west = [j for j in range(1, n - 1)]
    bdry = south + east + north + west
    bset = set(bdry)
    interior = [k for k in range(n * n) if k not in bset]
    return np.array(bdry + interior, dtype=np.int64)


def bounds_for_quad_subdivision(bounds: np.ndarray) -> np.ndarray:
    """Children of [xmin, xmax, ymin, ymax] in order SW, SE, NE, NW."""
    xmin, xmax, ymin, ymax = bounds
    xm = (xmin + xmax) / 2
    ym = (ymin + ymax) / 2
    return np.array(
        [
            [xmin, xm, ymin, ym],
            [xm, xmax, ymin, ym],
            [xm, xmax, ym, ymax],
            [xmin, xm, ym, ymax],
        ]
    )


def _subdivide_2D(bounds: np.ndarray, L: int) -> np.ndarray:
    b = np.asarray(bounds, dtype=np.float64).reshape(-1, 4)
    for _ in range(L):
        b = np.concatenate([bounds_for_quad_subdivision(x) for x in b], axis=0)
    return b


def bounds_to_cheby_points_2D(bounds: np.ndarray, cheby_pts_1d: np.ndarray) -> np.ndarray:
    n = cheby_pts_1d.shape[0]
    x = affine_transform(cheby_pts_1d, bounds[:2])
    y = affine_transform(cheby_pts_1d, bounds[2:])
    X, Y = np.meshgrid(x, y[::-1], indexing="ij")
    pts = np.stack([X.reshape(-1), Y.reshape(-1)], axis=-1)
    return pts[rearrange_indices_ext_int_2D(n)]


def compute_interior_Chebyshev_points_uniform_2D(root, L: int, p: int) -> np.ndarray:
    leaves = _subdivide_2D([root.xmin, root.xmax, root.ymin, root.ymax], L)
    c = chebyshev_points(p)
    return np.stack([bounds_to_cheby_points_2D(b, c) for b in leaves])


def compute_boundary_Gauss_points_uniform_2D(root, L: int, q: int) -> np.ndarray:
    g = gauss_points(q)
    n = 2**L
    xb = np.linspace(root.xmin, root.xmax, n + 1)
    yb = np.linspace(root.ymin, root.ymax, n + 1)
    xg = np.concatenate([affine_transform(g, xb[i : i + 2]) for i in range(n)])
    yg = np.concatenate([affine_transform(g, yb[i : i + 2]) for i in range(n)])
    return np.concatenate(
        [
            np.stack([xg, np.full_like(xg, root.ymin)], -1),
            np.stack([np.full_like(yg, root.xmax), yg], -1),
            np.stack([xg[::-1], np.full_like(xg, root.ymax)], -1),
            np.stack([np.full_like(yg, root.xmin), yg[::-1]], -1),
        ]
    )


# --------------------------------------------------------------------------
# 3D
# --------------------------------------------------------------------------
def rearrange_indices_ext_int_3D(p: int) -> np.ndarray:
    """Permutation of the ij-flattened p^3 grid (flat = i*p^2 + j*p + k for
    (x_i, y_j, z_k)) listing the faces x-, x+, then y-, y+ (minus points
    already listed), then z-, z+, then the interior
    (``_grid_creation_3D.py:421-458``)."""
    idx = np.arange(p**3)
    i = idx // (p * p)
    j = (idx // p) % p
    k = idx % p
    x_lo, x_hi = i == 0, i == p - 1
    used = x_lo | x_hi
    y_lo = (j == 0) & ~used
    y_hi = (j == p - 1) & ~used
    used = used | y_lo | y_hi
    z_lo = (k == 0) & ~used
    z_hi = (k == p - 1) & ~used
    used = used | z_lo | z_hi
    return np.concatenate(
        [idx[x_lo], idx[x_hi], idx[y_lo], idx[y_hi], idx[z_lo], idx[z_hi], idx[~used]]
    )


def bounds_for_oct_subdivision(bounds: np.ndarray) -> np.ndarray:
    """Children a..d (upper z: SW, SE, NE, NW in x-y), then e..h (lower z)."""
    x0, x1, y0, y1, z0, z1 = bounds
    xm, ym, zm = (x0 + x1) / 2, (y0 + y1) / 2, (z0 + z1) / 2
    quads = [[x0, xm, y0, ym], [xm, x1, y0, ym], [xm, x1, ym, y1], [x0, xm, ym, y1]]
    upper = [q + [zm, z1] for q in quads]
    lower = [q + [z0, zm] for q in quads]
    return np.array(upper + lower, dtype=np.float64)


def _subdivide_3D(bounds: np.ndarray, L: int) -> np.ndarray:
    b = np.asarray(bounds, dtype=np.float64).reshape(-1, 6)
    for _ in range(L):
        b = np.concatenate([bounds_for_oct_subdivision(x) for x in b], axis=0)
    return b


def bounds_to_cheby_points_3D(bounds: np.ndarray, cheby_pts_1d: np.ndarray) -> np.ndarray:
    p = cheby_pts_1d.shape[0]
    x = affine_transform(cheby_pts_1d, bounds[0:2])
    y = affine_transform(cheby_pts_1d, bounds[2:4])
    z = affine_transform(cheby_pts_1d, bounds[4:6])
    X, Y, Z = np.meshgrid(x, y, z, indexing="ij")
    pts = np.stack([X.reshape(-1), Y.reshape(-1), Z.reshape(-1)], -1)
    return pts[rearrange_indices_ext_int_3D(p)]


def compute_interior_Chebyshev_points_uniform_3D(root, L: int, p: int) -> np.ndarray:
    leaves = _subdivide_3D(
        [root.xmin, root.xmax, root.ymin, root.ymax, root.zmin, root.zmax], L
    )
    c = chebyshev_points(p)
    return np.stack([bounds_to_cheby_points_3D(b, c) for b in leaves])


def bounds_to_gauss_face(bounds: np.ndarray, gauss_pts_1d: np.ndarray) -> np.ndarray:
    """q x q Gauss grid on a face patch, meshgrid(indexing="ij") order."""
    a = affine_transform(gauss_pts_1d, bounds[0:2])
    b = affine_transform(gauss_pts_1d, bounds[2:4])
    A, B = np.meshgrid(a, b, indexing="ij")
    return np.stack([A.reshape(-1), B.reshape(-1)], -1)


def compute_boundary_Gauss_points_uniform_3D(root, L: int, q: int) -> np.ndarray:
    """Root boundary Gauss points, faces 1..6, patches in quad-subdivision
    order of each face's 2D bounds (``_grid_creation_3D.py:119-190``)."""
    g = gauss_points(q)
    xy = _subdivide_2D([root.xmin, root.xmax, root.ymin, root.ymax], L)
    yz = _subdivide_2D([root.ymin, root.ymax, root.zmin, root.zmax], L)
    xz = _subdivide_2D([root.xmin, root.xmax, root.zmin, root.zmax], L)
    gxy = np.concatenate([bounds_to_gauss_face(b, g) for b in xy])
    gyz = np.concatenate([bounds_to_gauss_face(b, g) for b in yz])
    gxz = np.concatenate([bounds_to_gauss_face(b, g) for b in xz])
    ones = np.ones(gyz.shape[0])
    f1 = np.stack([root.xmin * ones, gyz[:, 0], gyz[:, 1]], -1)
    f2 = np.stack([root.xmax * ones, gyz[:, 0], gyz[:, 1]], -1)
    f3 = np.stack([gxz[:, 0], root.ymin * ones, gxz[:, 1]], -1)
    f4 = np.stack([gxz[:, 0], root.ymax * ones, gxz[:, 1]], -1)
    f5 = np.stack([gxy[:, 0], gxy[:, 1], root.zmin * ones], -1)
    f6 = np.stack([gxy[:, 0], gxy[:, 1], root.zmax * ones], -1)
    return np.concatenate([f1, f2, f3, f4, f5, f6])


def get_all_uniform_leaves_2D(root, L: int):
    from ._discretization_tree import DiscretizationNode2D

    return [
        DiscretizationNode2D(xmin=b[0], xmax=b[1], ymin=b[2], ymax=b[3])
        for b in _subdivide_2D([root.xmin, root.xmax, root.ymin, root.ymax], L)
    ]


def get_all_uniform_leaves_3D(root, L: int):
    from ._discretization_tree import DiscretizationNode3D

    return [
        DiscretizationNode3D(xmin=b[0], xmax=b[1], ymin=b[2], ymax=b[3], zmin=b[4], zmax=b[5])
        for b in _subdivide_3D(
            [root.xmin, root.xmax, root.ymin, root.ymax, root.zmin, root.zmax], L
        )
    ]
