"""ORACLE (test infrastructure only) -- HPS grid <-> rectangular grid.

Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` may import this package; it is the checker, never the product.

Restates ``Domain.interp_to_interior_points`` / ``interp_from_interior_points``
(``_domain.py:99-294`` over ``_interpolation_methods.py:24-340``) point by point
from coordinates: every HPS point's value is the tensor-product Lagrange
interpolant of the samples (explicit products, not the barycentric formula
with its EPS substitution), and every evaluation point takes the Lagrange
interpolant of the first leaf, in leaf order, whose closed box contains it.
Pinned by exactness on polynomials of degree below the node counts (tests).
"""

from __future__ import annotations

import numpy as np

from .spectral import cheb_points, lagrange_matrix, leaf_points_2D, uniform_leaves_2D
from .hps3d import leaf_points_3D, uniform_leaves_3D


def interp_to_interior_2D(bounds, L, p, values, x, y):
    """values (n_x, n_y) on meshgrid(x, y, indexing="ij") -> (n_leaves, p^2)
    (``_interpolation_methods.py:276-300``)."""
    pts = leaf_points_2D(bounds, L, p)  # (n_leaves, p^2, 2)
    flat = pts.reshape(-1, 2)
    Lx = lagrange_matrix(np.asarray(x, float), flat[:, 0])  # (N, n_x)
    Ly = lagrange_matrix(np.asarray(y, float), flat[:, 1])  # (N, n_y)
    out = np.einsum("ni,ij,nj->n", Lx, values, Ly)
    return out.reshape(pts.shape[:2])


def interp_to_interior_3D(bounds, L, p, values, x, y, z):
    """values (n_x, n_y, n_z) -> (n_leaves, p^3) (``_interpolation_methods.py:309-340``)."""
    pts = leaf_points_3D(bounds, L, p)
    flat = pts.reshape(-1, 3)
    Lx = lagrange_matrix(np.asarray(x, float), flat[:, 0])
    Ly = lagrange_matrix(np.asarray(y, float), flat[:, 1])
    Lz = lagrange_matrix(np.asarray(z, float), flat[:, 2])
    out = np.einsum("ni,ijk,nj,nk->n", Lx, values, Ly, Lz)
    return out.reshape(pts.shape[:2])


def _first_containing(boxes, pt):
    for i, b in enumerate(boxes):
        d = len(pt)
        if all(b[2 * a] <= pt[a] <= b[2 * a + 1] for a in range(d)):
            return i
    return 0


def _leaf_interp_row(box, p, pt, order_pts):
    """Interpolation weights of the leaf's Chebyshev grid at pt, in the leaf's
    point order: the tensor Lagrange basis at the point's coordinates."""
    d = len(pt)
    c = cheb_points(p)
    w = np.ones(len(order_pts))
    for a in range(d):
        lo, hi = box[2 * a], box[2 * a + 1]
        nodes = 0.5 * (hi - lo) * c + 0.5 * (hi + lo)
        row = lagrange_matrix(nodes, np.array([pt[a]]))[0]  # weight of node j at pt
        # which 1D node each leaf point sits on
        idx = np.argmin(np.abs(order_pts[:, a, None] - nodes[None, :]), axis=1)
        w *= row[idx]
    return w


def to_points(pts, values, axes):
    """values on meshgrid(*axes, indexing="ij") -> every point of pts
    (n_leaves, n, d): tensor-product Lagrange interpolation."""
    flat = pts.reshape(-1, pts.shape[-1])
    Ls = [lagrange_matrix(np.asarray(a, float), flat[:, k]) for k, a in enumerate(axes)]
    if len(axes) == 2:
        out = np.einsum("ni,ij,nj->n", Ls[0], values, Ls[1])
    else:
        out = np.einsum("ni,ijk,nj,nk->n", Ls[0], values, Ls[1], Ls[2])
    return out.reshape(pts.shape[:2])


def from_leaves(boxes, pts, samples, axes):
    """HPS samples on leaves with the given boxes (leaf order) and points ->
    the points of meshgrid(*axes) (default "xy" order), flattened, with the
    first leaf holding each point (any tree)."""
    p = round(pts.shape[1] ** (1.0 / len(axes)))
    grids = np.meshgrid(*axes)
    flat = np.stack([g.ravel() for g in grids], -1)
    out = []
    for pt in flat:
        i = _first_containing(boxes, pt)
        w = _leaf_interp_row(boxes[i], p, pt, pts[i])
        out.append(np.tensordot(w, samples[i], axes=(0, 0)))
    return np.asarray(out)


def interp_from_interior_2D(bounds, L, p, samples, x_vals, y_vals):
    """samples (n_leaves, p^2[, s]) -> (vals reshaped (n_x, n_y[, s]), target
    points (n_y, n_x, 2)) with the reference's meshgrid(x, y) ("xy") point
    order (``_interpolation_methods.py:24-89``)."""
    boxes = uniform_leaves_2D(bounds, L)
    pts = leaf_points_2D(bounds, L, p)
    X, Y = np.meshgrid(x_vals, y_vals)
    target = np.stack([X, Y], -1)
    flat = target.reshape(-1, 2)
    out = []
    for pt in flat:
        i = _first_containing(boxes, pt)
        w = _leaf_interp_row(boxes[i], p, pt, pts[i])
        out.append(np.tensordot(w, samples[i], axes=(0, 0)))
    out = np.asarray(out)
    nx, ny = len(x_vals), len(y_vals)
    return out.reshape((nx, ny) + samples.shape[2:]), target


def interp_from_interior_3D(bounds, L, p, samples, x_vals, y_vals, z_vals):
    """(``_interpolation_methods.py:153-219``): meshgrid(x, y, z) ("xy") order,
    reshaped to (n_x, n_y, n_z)."""
    boxes = uniform_leaves_3D(bounds, L)
    pts = leaf_points_3D(bounds, L, p)
    X, Y, Z = np.meshgrid(x_vals, y_vals, z_vals)
    target = np.stack([X, Y, Z], -1)
    flat = np.stack([X.ravel(), Y.ravel(), Z.ravel()], -1)
    out = []
    for pt in flat:
        i = _first_containing(boxes, pt)
        w = _leaf_interp_row(boxes[i], p, pt, pts[i])
        out.append(w @ samples[i])
    out = np.asarray(out)
    return out.reshape(len(x_vals), len(y_vals), len(z_vals)), target
