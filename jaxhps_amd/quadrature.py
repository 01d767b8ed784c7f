"""Spectral primitives on the host (float64 numpy).

Same conventions and call signatures as the reference's ``quadrature``
package (``quadrature/__init__.py:1-26``).  These matrices are built once per
problem on the host and copied to HBM; nothing here is on the hot path.
"""

from __future__ import annotations

import numpy as np

_EPS = np.finfo(np.float64).eps


def chebyshev_points(n: int) -> np.ndarray:
    """Second-kind Chebyshev points on [-1, 1], ascending.

    out[i] = cos(pi * (n-1-i) / (n-1)); follows
    ``quadrature/_discretization.py:13-37``.
    """
    angles = np.pi * (np.arange(n - 1, -1, -1, dtype=np.float64) / (n - 1))
    return np.cos(angles)


def chebyshev_weights(n: int, bounds) -> np.ndarray:
    """Clenshaw-Curtis weights on [a, b] (``quadrature/_discretization.py:41-86``).

    Computed from the closed-form cosine sums rather than the FFT mirror trick;
    both give the same rule (the tests compare against exact integrals).
    """
    a, b = bounds
    N = n - 1
    theta = np.pi * np.arange(n) / N
    w = np.zeros(n)
    for j in range(n):
        s = 0.0
        for k in range(1, N // 2 + 1):
            bk = 1.0 if 2 * k == N else 2.0
            s += bk / (4 * k * k - 1) * np.cos(2 * k * theta[j])
        cj = 1.0 if (j == 0 or j == N) else 2.0
        w[j] = cj / N * (1.0 - s)
    # ascending order of points (theta decreasing) -- the rule is symmetric
    return w[::-1] * (b - a) / 2.0


def affine_transform(pts: np.ndarray, ab) -> np.ndarray:
    """Map points on [-1, 1] to [a, b] (``quadrature/_discretization.py:89-102``)."""
    a, b = ab
    return 0.5 * (b - a) * np.asarray(pts) + 0.5 * (a + b)


def gauss_points(n: int) -> np.ndarray:
    """Gauss-Legendre nodes on [-1, 1] (``quadrature/_discretization.py:105-116``)."""
    return np.polynomial.legendre.leggauss(n)[0]


def differentiation_matrix_1D(points: np.ndarray) -> np.ndarray:
    """Chebyshev spectral differentiation matrix on the given nodes.

    Trefethen, *Spectral Methods in MATLAB* ch. 6 (``cheb.m``), in the
    ascending node order of ``chebyshev_points``; follows
    ``quadrature/_differentiation.py:9-47`` (off-diagonal c_i/c_j/(x_i-x_j),
    diagonal = minus the row sum).
    """
    points = np.asarray(points, dtype=np.float64)
    p = points.shape[0]
    c = np.ones(p)
    c[0] = 2.0
    c[-1] = 2.0
    c[1::2] *= -1.0
    dx = points[:, None] - points[None, :]
    d = np.outer(c, 1.0 / c) / (dx + np.eye(p))
    d = d - np.diag(d.sum(axis=1))
    return d


def _bary_weights(x: np.ndarray) -> np.ndarray:
    diff = x[:, None] - x[None, :]
    np.fill_diagonal(diff, 1.0)
    return np.prod(diff, axis=0)


def barycentric_lagrange_interpolation_matrix_1D(from_pts, to_pts) -> np.ndarray:
    """Barycentric Lagrange interpolation matrix, shape (len(to), len(from)).

    Rows whose target coincides with a source node become unit rows
    (``quadrature/_interpolation.py:18-114``).
    """
    from_pts = np.asarray(from_pts, dtype=np.float64)
    to_pts = np.asarray(to_pts, dtype=np.float64)
    w = _bary_weights(from_pts)
    with np.errstate(divide="ignore", invalid="ignore"):
        dist = to_pts[:, None] - from_pts[None, :]
        norm = np.sum(1.0 / (w[None, :] * dist), axis=1)
        mat = 1.0 / (dist * w[None, :] * norm[:, None])
    matches = to_pts[:, None] == from_pts[None, :]
    has = matches.any(axis=1)
    mat = np.where(has[:, None], matches.astype(np.float64), mat)
    return mat


def _bary_factor_eps(from_pts, to_pts):
    """1D barycentric factor with the reference's EPS replacement for exact
    coincidences (``quadrature/_interpolation.py:184-191``)."""
    w = _bary_weights(from_pts)
    dist = to_pts[None, :] - from_pts[:, None]  # (n_from, n_to)
    dist = np.where(dist == 0, _EPS, dist)
    norm = np.sum(1.0 / (w[:, None] * dist), axis=0)  # (n_to,)
    # f[i, k] = 1 / (dist[k, i] * w[k] * norm[i])
    return 1.0 / (dist.T * w[None, :] * norm[:, None])


def barycentric_lagrange_interpolation_matrix_2D(
    from_pts_x, from_pts_y, to_pts_x, to_pts_y
) -> np.ndarray:
    """Tensor-product barycentric interpolation, meshgrid(indexing="ij") order
    on both sides (``quadrature/_interpolation.py:118-213``)."""
    fx = _bary_factor_eps(np.asarray(from_pts_x, float), np.asarray(to_pts_x, float))
    fy = _bary_factor_eps(np.asarray(from_pts_y, float), np.asarray(to_pts_y, float))
    return np.kron(fx, fy)


def barycentric_lagrange_interpolation_matrix_3D(
    from_pts_x, from_pts_y, from_pts_z, to_pts_x, to_pts_y, to_pts_z
) -> np.ndarray:
    """3D tensor-product version (``quadrature/_interpolation.py:217-326``)."""
    fx = _bary_factor_eps(np.asarray(from_pts_x, float), np.asarray(to_pts_x, float))
    fy = _bary_factor_eps(np.asarray(from_pts_y, float), np.asarray(to_pts_y, float))
    fz = _bary_factor_eps(np.asarray(from_pts_z, float), np.asarray(to_pts_z, float))
    return np.kron(fx, np.kron(fy, fz))


def meshgrid_to_lst_of_pts(X: np.ndarray, Y: np.ndarray) -> np.ndarray:
    """Stack two (n, n) meshgrid arrays into (n*n, 2) points."""
    return np.stack([X.reshape(-1), Y.reshape(-1)], axis=-1)


__all__ = [
    "differentiation_matrix_1D",
    "chebyshev_points",
    "chebyshev_weights",
    "affine_transform",
    "gauss_points",
    "barycentric_lagrange_interpolation_matrix_1D",
    "barycentric_lagrange_interpolation_matrix_2D",
    "barycentric_lagrange_interpolation_matrix_3D",
    "meshgrid_to_lst_of_pts",
]
