"""The adaptive 3D DtN oracle pinned by the uniform 3D restatement, exact
polynomial solutions on non-uniform octrees and exact projections."""

import numpy as np
import pytest

from oracle import adaptive3d as ad
from oracle import hps2d, hps3d

B0 = (0.0, 1.0, 0.0, 1.0, 0.0, 1.0)


def _uniform(L):
    root = ad.Node(B0)
    todo = [root]
    for _ in range(L):
        nxt = []
        for nd in todo:
            ad.split(nd)
            nxt += nd.children
        todo = nxt
    return root


def _nonuniform():
    root = ad.Node(B0)
    ad.split(root)
    ad.split(root.children[0])
    ad.split(root.children[6])
    return root


def test_projection_ops_exact():
    q = 4
    L41, L14 = ad.projection_ops(q)
    g = np.polynomial.legendre.leggauss(q)[0]
    f = lambda a, b: a**3 - 2 * a * b**2 + b + 1
    A, Bm = np.meshgrid(g, g, indexing="ij")
    coarse = f(A, Bm).ravel()
    fine = []
    for ca, cb in ((-0.5, -0.5), (0.5, -0.5), (0.5, 0.5), (-0.5, 0.5)):
        fa, fb = np.meshgrid(ca + 0.5 * g, cb + 0.5 * g, indexing="ij")
        fine.append(f(fa, fb).ravel())
    fine = np.concatenate(fine)
    assert np.abs(L41 @ coarse - fine).max() < 1e-13
    assert np.abs(L14 @ fine - coarse).max() < 1e-13


def test_uniform_octree_matches_uniform_restatement():
    p, q, L = 5, 2, 1
    root = _uniform(L)
    rng = np.random.default_rng(0)
    nl = 8**L
    coeffs = {"D_xx": 1 + 0.2 * rng.random((nl, p**3)), "D_yy": np.ones((nl, p**3)), "D_zz": np.ones((nl, p**3)),
              "I": rng.standard_normal((nl, p**3))}
    f = rng.standard_normal((nl, p**3, 2))
    nb = ad.boundary_points(root, q).shape[0]
    assert np.allclose(ad.boundary_points(root, q), hps3d.boundary_gauss_points_3D(B0, L, q), atol=1e-15)
    assert np.allclose(ad.interior_points(root, p), hps3d.leaf_points_3D(B0, L, p), atol=1e-15)
    bdry = rng.standard_normal((nb, 2))
    u, T = ad.solve(root, p, q, coeffs, f, bdry)
    ops = hps3d.leaf_ops_3D(p, q, 0.5 / 2**L)
    order = ("D_xx", "D_xy", "D_yy", "D_xz", "D_yz", "D_zz", "D_x", "D_y", "D_z", "I")
    A = hps2d.assemble_diff_operator(coeffs, ops, order=order)
    Y, Tl, v, h = hps2d.leaf_dtn(A, f, ops["P"], ops["Q"])
    S_lst, g_lst, Tu, _ = hps3d.merge_stage_3d(Tl, h, L)
    uu, _ = hps3d.down_pass_3d(bdry, S_lst, g_lst, Y, v)
    assert np.abs(T - Tu).max() < 1e-10 * np.abs(Tu).max()
    assert np.abs(u - uu).max() < 1e-10 * np.abs(uu).max()


def test_compression_lists_and_counts():
    root = _nonuniform()
    q = 2
    assert ad.face_counts(root, q) == tuple(q * q * k for k in (7, 7, 7, 7, 7, 7))
    fl = ad.compression_lists(root)
    # interface 9 (a x+ | b x-): a refined -> its 4 face leaves compress
    assert fl[0] == ([True] * 4, [False])
    # interface 17 (a z- | e z+): a refined on z-
    assert fl[8] == ([True] * 4, [False])
    # interface 14 (f y+ | g y-): g refined
    assert fl[5] == ([False], [True] * 4)
    assert ad.level_restricted(root)


@pytest.mark.parametrize("p,q", [(6, 4)])
def test_polynomial_solution_exact_on_non_uniform_octree(p, q):
    root = _nonuniform()
    pts = ad.interior_points(root, p)
    X, Y, Z = pts[..., 0], pts[..., 1], pts[..., 2]
    u = X**3 - 2 * X * Y**2 + Z**3 + X * Y * Z
    a = 1 + X**2
    uxx, uyy, uzz = 6 * X, -4 * X, 6 * Z
    c = np.cos(Y)
    coeffs = {"D_xx": a, "D_yy": np.ones_like(X), "D_zz": np.ones_like(X), "I": c}
    f = a * uxx + uyy + uzz + c * u
    bp = ad.boundary_points(root, q)
    g = bp[:, 0] ** 3 - 2 * bp[:, 0] * bp[:, 1] ** 2 + bp[:, 2] ** 3 + bp[:, 0] * bp[:, 1] * bp[:, 2]
    sol, T = ad.solve(root, p, q, coeffs, f, g)
    assert T.shape == (bp.shape[0], bp.shape[0])
    assert np.abs(sol[..., 0] - u).max() < 1e-9
