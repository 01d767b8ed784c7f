"""The interpolation oracle (HPS grid <-> rectangular grid) pinned by
exactness on polynomials and by the first-leaf rule on shared edges."""

import numpy as np
import pytest

from oracle import interp as oi
from oracle.hps3d import leaf_points_3D
from oracle.spectral import leaf_points_2D


def _f2(X, Y):
    return X**5 - 2 * X**2 * Y**3 + Y**4 + 0.3


def test_to_interior_2D_exact_on_polynomials():
    b = (-1.0, 1.0, -0.5, 0.5)
    x = np.linspace(-1, 1, 6)
    y = np.linspace(-0.5, 0.5, 5)
    V = _f2(*np.meshgrid(x, y, indexing="ij"))
    out = oi.interp_to_interior_2D(b, 2, 5, V, x, y)
    pts = leaf_points_2D(b, 2, 5)
    assert np.abs(out - _f2(pts[..., 0], pts[..., 1])).max() < 1e-12


def test_to_interior_3D_exact_on_polynomials():
    b = (0.0, 1.0, 0.0, 2.0, -1.0, 1.0)
    x, y, z = np.linspace(0, 1, 4), np.linspace(0, 2, 3), np.linspace(-1, 1, 5)
    f = lambda X, Y, Z: X**3 * Z - Y**2 + Z**4 * X
    V = f(*np.meshgrid(x, y, z, indexing="ij"))
    out = oi.interp_to_interior_3D(b, 1, 4, V, x, y, z)
    pts = leaf_points_3D(b, 1, 4)
    assert np.abs(out - f(pts[..., 0], pts[..., 1], pts[..., 2])).max() < 1e-12


def test_from_interior_2D_exact_and_layout():
    b = (-1.0, 1.0, -0.5, 0.5)
    g = lambda X, Y: X**3 - X * Y**2 + 2 * Y**3 - 1
    pts = leaf_points_2D(b, 1, 4)
    S = g(pts[..., 0], pts[..., 1])
    xv, yv = np.linspace(-1, 1, 7), np.linspace(-0.5, 0.5, 4)
    vals, tgt = oi.interp_from_interior_2D(b, 1, 4, S, xv, yv)
    assert vals.shape == (7, 4) and tgt.shape == (4, 7, 2)
    # the reference flattens meshgrid(x, y) ("xy") points and reshapes to (n_x, n_y)
    assert np.abs(vals - g(tgt[..., 0], tgt[..., 1]).reshape(7, 4)).max() < 1e-12


def test_from_interior_first_leaf_on_shared_edges():
    """A leaf-wise constant field: points on shared edges take the first leaf
    in leaf order (SW, SE, NE, NW at L = 1)."""
    b = (0.0, 2.0, 0.0, 2.0)
    S = np.repeat(np.arange(4.0)[:, None], 16, axis=1)
    vals, tgt = oi.interp_from_interior_2D(b, 1, 4, S, np.array([1.0]), np.array([0.5, 1.0, 1.5]))
    # (1, 0.5): SW|SE -> SW (0); (1, 1): centre -> SW (0); (1, 1.5): NE|NW -> NE (2)
    assert list(vals.reshape(-1)) == pytest.approx([0.0, 0.0, 2.0])


def test_from_interior_3D_exact():
    b = (0.0, 1.0, 0.0, 1.0, 0.0, 1.0)
    f = lambda X, Y, Z: X**2 * Y - Z**3 + X * Y * Z
    pts = leaf_points_3D(b, 1, 4)
    S = f(pts[..., 0], pts[..., 1], pts[..., 2])
    xv, yv, zv = np.linspace(0, 1, 3), np.linspace(0, 1, 2), np.linspace(0, 1, 4)
    vals, tgt = oi.interp_from_interior_3D(b, 1, 4, S, xv, yv, zv)
    flat = np.stack([tgt[..., 0].ravel(), tgt[..., 1].ravel(), tgt[..., 2].ravel()], -1)
    assert np.abs(vals.reshape(-1) - f(flat[:, 0], flat[:, 1], flat[:, 2])).max() < 1e-12
