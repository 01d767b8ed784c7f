"""GPU parity of the Domain interpolation utilities (HPS grid <-> rectangular
grid) against the oracle, plus exactness on polynomials of low degree."""

import numpy as np
import pytest
import torch

from oracle import interp as oi
from oracle.hps3d import leaf_points_3D
from oracle.spectral import leaf_points_2D

pytestmark = pytest.mark.gpu


def _np(t):
    return t.detach().cpu().numpy()


def _dom2(p, L, b=(-1.0, 1.0, -0.5, 0.5)):
    import jaxhps_amd as jh

    return jh.Domain(p=p, q=p - 2, root=jh.DiscretizationNode2D(*b), L=L), b


def _dom3(p, L, b=(0.0, 1.0, -1.0, 1.0, 0.0, 2.0)):
    import jaxhps_amd as jh

    return jh.Domain(p=p, q=p - 2, root=jh.DiscretizationNode3D(*b), L=L), b


@pytest.mark.parametrize("p,L,cplx", [(6, 2, False), (16, 3, True), (8, 4, False)])
def test_to_interior_2D(p, L, cplx):
    dom, b = _dom2(p, L)
    rng = np.random.default_rng(p + L)
    x = np.sort(rng.uniform(b[0], b[1], 9))
    y = np.sort(rng.uniform(b[2], b[3], 7))
    V = rng.standard_normal((9, 7)) + (1j * rng.standard_normal((9, 7)) if cplx else 0)
    got = _np(dom.interp_to_interior_points(torch.as_tensor(V), x, y))
    want = oi.interp_to_interior_2D(b, L, p, V, x, y)
    assert got.shape == (4**L, p * p)
    assert np.abs(got - want).max() <= 1e-9 * max(np.abs(want).max(), 1.0)
    # a polynomial of degree below the sample counts comes back exactly
    f = lambda X, Y: X**4 - 3 * X * Y**3 + Y**2 - 0.5
    xs, ys = np.linspace(b[0], b[1], 6), np.linspace(b[2], b[3], 5)
    got = _np(dom.interp_to_interior_points(f(*np.meshgrid(xs, ys, indexing="ij")), xs, ys))
    pts = leaf_points_2D(b, L, p)
    assert np.abs(got - f(pts[..., 0], pts[..., 1])).max() < 1e-11


@pytest.mark.parametrize("p,L,cplx", [(4, 1, False), (8, 2, True)])
def test_to_interior_3D(p, L, cplx):
    dom, b = _dom3(p, L)
    rng = np.random.default_rng(3 * p + L)
    x = np.sort(rng.uniform(b[0], b[1], 5))
    y = np.sort(rng.uniform(b[2], b[3], 4))
    z = np.sort(rng.uniform(b[4], b[5], 6))
    V = rng.standard_normal((5, 4, 6)) + (1j * rng.standard_normal((5, 4, 6)) if cplx else 0)
    got = _np(dom.interp_to_interior_points(torch.as_tensor(V), x, y, z))
    want = oi.interp_to_interior_3D(b, L, p, V, x, y, z)
    assert got.shape == (8**L, p**3)
    assert np.abs(got - want).max() <= 1e-9 * max(np.abs(want).max(), 1.0)


@pytest.mark.parametrize("p,L,nsrc,cplx", [(6, 2, None, False), (16, 3, 3, True), (10, 4, None, True)])
def test_from_interior_2D(p, L, nsrc, cplx):
    dom, b = _dom2(p, L)
    rng = np.random.default_rng(p * L)
    shape = (4**L, p * p) + ((nsrc,) if nsrc else ())
    S = rng.standard_normal(shape) + (1j * rng.standard_normal(shape) if cplx else 0)
    # include the box edges and leaf edges, where the first-leaf rule decides
    xv = np.concatenate([np.linspace(b[0], b[1], 2**L + 1), rng.uniform(b[0], b[1], 6)])
    yv = np.concatenate([np.linspace(b[2], b[3], 2**L + 1), rng.uniform(b[2], b[3], 5)])
    vals, tgt = dom.interp_from_interior_points(torch.as_tensor(S), xv, yv)
    want, wtgt = oi.interp_from_interior_2D(b, L, p, S, xv, yv)
    assert tuple(vals.shape) == want.shape
    assert np.array_equal(np.asarray(tgt), wtgt)
    assert np.abs(_np(vals) - want).max() <= 1e-9 * max(np.abs(S).max(), 1.0)


def test_from_interior_2D_polynomial_exact():
    p, L = 8, 2
    dom, b = _dom2(p, L)
    g = lambda X, Y: X**7 - 2 * X**3 * Y**4 + Y**7 + 1
    pts = leaf_points_2D(b, L, p)
    xv, yv = np.linspace(b[0], b[1], 13), np.linspace(b[2], b[3], 11)
    vals, tgt = dom.interp_from_interior_points(torch.as_tensor(g(pts[..., 0], pts[..., 1])), xv, yv)
    exact = g(tgt[..., 0], tgt[..., 1]).reshape(13, 11)
    assert np.abs(_np(vals) - exact).max() < 1e-11


@pytest.mark.parametrize("p,L,cplx", [(4, 1, False), (6, 2, True)])
def test_from_interior_3D(p, L, cplx):
    dom, b = _dom3(p, L)
    rng = np.random.default_rng(7 * p + L)
    shape = (8**L, p**3)
    S = rng.standard_normal(shape) + (1j * rng.standard_normal(shape) if cplx else 0)
    xv = np.concatenate([np.linspace(b[0], b[1], 2**L + 1), rng.uniform(b[0], b[1], 2)])
    yv = np.concatenate([np.linspace(b[2], b[3], 2**L + 1), rng.uniform(b[2], b[3], 2)])
    zv = np.concatenate([np.linspace(b[4], b[5], 2**L + 1), rng.uniform(b[4], b[5], 3)])
    vals, tgt = dom.interp_from_interior_points(torch.as_tensor(S), xv, yv, zv)
    want, wtgt = oi.interp_from_interior_3D(b, L, p, S, xv, yv, zv)
    assert tuple(vals.shape) == want.shape
    assert np.array_equal(np.asarray(tgt), wtgt)
    assert np.abs(_np(vals) - want).max() <= 1e-9 * max(np.abs(S).max(), 1.0)


def test_from_interior_3D_polynomial_exact():
    p, L = 6, 1
    dom, b = _dom3(p, L)
    f = lambda X, Y, Z: X**5 * Y - Z**4 * X + Y**3 * Z**2
    pts = leaf_points_3D(b, L, p)
    xv, yv, zv = np.linspace(b[0], b[1], 5), np.linspace(b[2], b[3], 4), np.linspace(b[4], b[5], 6)
    vals, tgt = dom.interp_from_interior_points(torch.as_tensor(f(pts[..., 0], pts[..., 1], pts[..., 2])), xv, yv, zv)
    exact = f(tgt[..., 0].ravel(), tgt[..., 1].ravel(), tgt[..., 2].ravel()).reshape(5, 4, 6)
    assert np.abs(_np(vals) - exact).max() < 1e-11


def test_round_trip_grid_to_hps_and_back():
    """rectangular grid -> HPS grid -> same rectangular grid is the identity
    for polynomials the first map reproduces."""
    p, L = 10, 2
    dom, b = _dom2(p, L)
    f = lambda X, Y: np.cos(0) * (X**3 + X * Y - Y**2)
    xs, ys = np.linspace(b[0], b[1], 5), np.linspace(b[2], b[3], 4)
    hps = dom.interp_to_interior_points(f(*np.meshgrid(xs, ys, indexing="ij")), xs, ys)
    vals, tgt = dom.interp_from_interior_points(hps, xs, ys)
    assert np.abs(_np(vals) - f(tgt[..., 0], tgt[..., 1]).reshape(5, 4)).max() < 1e-11


def test_interp_on_adaptive_trees_2D_and_3D():
    """Both directions on non-uniform trees (the leaf of a point is the first
    adaptive leaf holding it)."""
    import jaxhps_amd as jh
    from jaxhps_amd import _adaptive as pa
    from jaxhps_amd import _adaptive3d as pa3

    rng = np.random.default_rng(11)
    root = jh.DiscretizationNode2D(-1.0, 1.0, -0.5, 0.5)
    pa.add_four_children(root, root=root, q=4)
    pa.add_four_children(root.children[1], root=root, q=4)
    pa.add_four_children(root.children[1].children[3], root=root, q=4)
    dom = jh.Domain(p=7, q=4, root=root)
    boxes = [(lf.xmin, lf.xmax, lf.ymin, lf.ymax) for lf in jh.get_all_leaves(root)]
    x, y = np.linspace(-1, 1, 6), np.linspace(-0.5, 0.5, 5)
    V = rng.standard_normal((6, 5))
    got = _np(dom.interp_to_interior_points(torch.as_tensor(V), x, y))
    assert np.abs(got - oi.to_points(dom.interior_points, V, (x, y))).max() < 1e-9 * np.abs(V).max() * 10
    S = rng.standard_normal(dom.interior_points.shape[:2])
    xv = np.concatenate([np.linspace(-1, 1, 9), rng.uniform(-1, 1, 4)])
    yv = np.concatenate([np.linspace(-0.5, 0.5, 9), rng.uniform(-0.5, 0.5, 3)])
    vals, _ = dom.interp_from_interior_points(torch.as_tensor(S), xv, yv)
    want = oi.from_leaves(boxes, dom.interior_points, S, (xv, yv)).reshape(xv.size, yv.size)
    assert np.abs(_np(vals) - want).max() <= 1e-9 * np.abs(S).max()

    root3 = jh.DiscretizationNode3D(0.0, 1.0, 0.0, 1.0, 0.0, 1.0)
    pa3.add_eight_children(root3, root=root3, q=2)
    pa3.add_eight_children(root3.children[2], root=root3, q=2)
    dom3 = jh.Domain(p=4, q=2, root=root3)
    boxes3 = [(lf.xmin, lf.xmax, lf.ymin, lf.ymax, lf.zmin, lf.zmax) for lf in jh.get_all_leaves(root3)]
    S3 = rng.standard_normal(dom3.interior_points.shape[:2])
    ax = [np.linspace(0, 1, 5), np.linspace(0, 1, 3), np.linspace(0, 1, 4)]
    vals3, _ = dom3.interp_from_interior_points(torch.as_tensor(S3), *ax)
    want3 = oi.from_leaves(boxes3, dom3.interior_points, S3, ax).reshape(5, 3, 4)
    assert np.abs(_np(vals3) - want3).max() <= 1e-9 * np.abs(S3).max()
    V3 = rng.standard_normal((3, 4, 2))
    ax3 = [np.linspace(0, 1, 3), np.linspace(0, 1, 4), np.linspace(0, 1, 2)]
    got3 = _np(dom3.interp_to_interior_points(torch.as_tensor(V3), *ax3))
    assert np.abs(got3 - oi.to_points(dom3.interior_points, V3, ax3)).max() < 1e-9 * 10
