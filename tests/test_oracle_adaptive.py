"""The adaptive 2D DtN oracle pinned by the uniform restatement, exact
polynomial solutions on non-uniform trees, the panel-norm closed forms of
the reference's tests and the level-restriction invariant."""

import numpy as np
import pytest

from oracle import adaptive2d as ad
from oracle import hps2d
from oracle import spectral as osp


def _uniform(L, b=(0.0, 1.0, 0.0, 1.0)):
    root = ad.Node(*b)
    todo = [root]
    for _ in range(L):
        nxt = []
        for nd in todo:
            ad.split(nd)
            nxt += nd.children
        todo = nxt
    return root


def _ref_tree():
    """The non-uniform tree of the reference's adaptive solve test."""
    root = ad.Node(0.0, 1.0, 0.0, 1.0)
    ad.split(root)
    ad.split(root.children[0])
    ad.split(root.children[1])
    ad.split(root.children[0].children[1])
    return root


def test_uniform_tree_matches_uniform_restatement():
    p, q, L = 6, 4, 2
    root = _uniform(L)
    rng = np.random.default_rng(0)
    n_leaves = 4**L
    coeffs = {"D_xx": 1 + 0.2 * rng.random((n_leaves, p * p)), "D_yy": np.ones((n_leaves, p * p)),
              "I": rng.standard_normal((n_leaves, p * p))}
    f = rng.standard_normal((n_leaves, p * p, 2))
    bdry = rng.standard_normal((4 * q * 2**L, 2))
    u, T = ad.solve(root, p, q, coeffs, f, bdry)
    ops = osp.leaf_ops_2D(p, q, 0.5 / 2**L)
    A = hps2d.assemble_diff_operator(coeffs, ops)
    Y, Tl, v, h = hps2d.leaf_dtn(A, f, ops["P"], ops["Q"])
    S_lst, g_lst, Tu, _ = hps2d.merge_stage(Tl, h, L, "dtn")
    uu, _ = hps2d.down_pass(bdry, S_lst, g_lst, Y, v, "dtn")
    assert np.abs(T - Tu).max() < 1e-10 * np.abs(Tu).max()
    assert np.abs(u - uu).max() < 1e-10 * np.abs(uu).max()
    assert np.allclose(ad.boundary_points(root, q), osp.boundary_gauss_points_2D((0, 1, 0, 1), L, q), atol=1e-15)
    assert np.allclose(ad.interior_points(root, p), osp.leaf_points_2D((0, 1, 0, 1), L, p), atol=1e-15)


def test_compression_lists_of_the_reference_tree():
    root = _ref_tree()
    a, b, c, d = root.children
    fl = ad.compression_lists(a, b, c, d)
    # interface 5 (a|b): a has SE.SW-level leaves below y = 0.25 only on its
    # SE child's east side: a5 = [T, T, F], b5 = [F, F]
    assert fl[0] == [True, True, False] and fl[1] == [False, False]
    # interface 6 (b|c): b's top leaves are level 2, c is one leaf
    assert fl[2] == [True, True] and fl[3] == [False]
    # interface 7 (c|d): both single leaves
    assert fl[4] == [False] and fl[5] == [False]
    # interface 8 (d|a): a's top leaves are level 2
    assert fl[6] == [False] and fl[7] == [True, True]
    assert ad.side_counts(root, 4) == (20, 12, 8, 12)
    assert ad.level_restricted(root)


@pytest.mark.parametrize("p,q", [(8, 6), (10, 8)])
def test_polynomial_solution_exact_on_non_uniform_tree(p, q):
    """u = x^3 - 2 x y^2 + y^3 + x with a u_xx + u_yy + c u = f: every map in
    the chain (leaf solves, compression, refinement of interface data) is
    exact for this degree."""
    root = _ref_tree()
    pts = ad.interior_points(root, p)
    X, Y = pts[..., 0], pts[..., 1]
    u = X**3 - 2 * X * Y**2 + Y**3 + X
    uxx, uyy = 6 * X, -4 * X + 6 * Y
    a, c = 1 + X**2, np.sin(Y)
    coeffs = {"D_xx": a, "D_yy": np.ones_like(X), "I": c}
    f = a * uxx + uyy + c * u
    bp = ad.boundary_points(root, q)
    g = bp[:, 0] ** 3 - 2 * bp[:, 0] * bp[:, 1] ** 2 + bp[:, 1] ** 3 + bp[:, 0]
    sol, T = ad.solve(root, p, q, coeffs, f, g)
    assert T.shape == (bp.shape[0], bp.shape[0])
    assert np.abs(sol[..., 0] - u).max() < 1e-10


def test_panel_norm_closed_forms():
    """The reference's test_adaptive_discretization_2D closed forms."""
    p = 4
    assert ad.panel_sq_l2(3 * np.ones(p * p), (0.0, 1.0, 0.0, 1.0), p) == pytest.approx(9.0)
    p = 16
    pts = ad._panel_pts((0.0, 1.0, 0.0, 1.0), p)
    assert ad.panel_sq_l2(pts[:, 1] + pts[:, 0] ** 2, (0.0, 1.0, 0.0, 1.0), p) == pytest.approx(13 / 15)
    b = (-np.pi / 2, np.pi / 2, -np.pi / 2, np.pi / 2)
    assert ad.panel_sq_l2(3 * np.ones(p * p), b, p) == pytest.approx(9 * np.pi**2)


@pytest.mark.parametrize("l2", [False, True])
def test_refinement_is_level_restricted_and_resolves(l2):
    p = 6
    root = ad.Node(-1.0, 1.0, -1.0, 1.0)
    f = lambda x: np.exp(-40 * ((x[..., 0] - 0.3) ** 2 + (x[..., 1] + 0.2) ** 2))
    ad.refine(root, f, 1e-4, p, restrict=True, l2=l2)
    lv = ad.leaves(root)
    assert len(lv) > 4 and ad.level_restricted(root)
    # refined near the bump, coarse far away
    depth_at = lambda x, y: [lf.depth for lf in lv if lf.xmin <= x <= lf.xmax and lf.ymin <= y <= lf.ymax][0]
    assert depth_at(0.3, -0.2) > depth_at(-0.95, 0.95)
    # the refined grid resolves f: interpolation to a finer grid changes it little
    R = ad._refine_operator(p)
    worst = max(np.abs(R @ f(ad._panel_pts(ad.bounds(lf), p)) - f(ad._children_pts(ad.bounds(lf), p)[0])).max()
                for lf in lv)
    assert worst < 1e-3


def test_refine_operator_exact_on_polynomials():
    p = 6
    R = ad._refine_operator(p)
    f = lambda x: x[..., 0] ** 5 - x[..., 0] * x[..., 1] ** 3 + 2
    b = (-1.0, 1.0, -1.0, 1.0)
    assert np.abs(R @ f(ad._panel_pts(b, p)) - f(ad._children_pts(b, p)[0])).max() < 1e-12
