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
