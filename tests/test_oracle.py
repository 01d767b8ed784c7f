"""The FP64 oracle against what the reference's own tests fix and against
what the mathematics fixes (CPU only).

* Textbook values: Chebyshev D-matrices for n = 2, 3, 4 (Trefethen, Spectral
  Methods in MATLAB, cheb.m; the reference's
  tests/test_quadrature/test_differentiation.py:11-67), Chebyshev points
  (test_discretization.py:103-115), the barycentric example
  (test_interpolation.py:14-24).
* Closed forms of tests/test_accuracy/cases.py at the reference's tolerances
  (polynomial cases atol 1e-12, Helmholtz-like atol 1e-8).
* The global discrete system assembled from coordinates (no index tables):
  the stage-by-stage solution must agree to rounding.
* No-source build + up pass == with-source build
  (test_nosource_accuracy.py:166-280).
"""

import numpy as np
import pytest

import cases2d as cs
from oracle import global_system as gs
from oracle import hps2d
from oracle import spectral as osp


# ---------------------------------------------------------------------------
# spectral primitives
# ---------------------------------------------------------------------------
def test_cheb_points_textbook():
    assert np.allclose(osp.cheb_points(2), [-1.0, 1.0], atol=1e-15)
    assert np.allclose(osp.cheb_points(3), [-1.0, 0.0, 1.0], atol=1e-15)


def test_diff_matrix_textbook():
    # cheb.m for N = 1, 2, 3 (n = 2, 3, 4 points), reordered to ascending nodes
    D2 = np.array([[-0.5, 0.5], [-0.5, 0.5]])
    D3 = np.array([[-1.5, 2.0, -0.5], [-0.5, 0.0, 0.5], [0.5, -2.0, 1.5]])
    x4 = osp.cheb_points(4)
    D4_ref = np.array(
        [
            [-19 / 6, 4.0, -4 / 3, 0.5],
            [-1.0, 1 / 3, 1.0, -1 / 3],
            [1 / 3, -1.0, -1 / 3, 1.0],
            [-0.5, 4 / 3, -4.0, 19 / 6],
        ]
    )
    assert np.allclose(osp.diff_matrix(osp.cheb_points(2)), D2, atol=1e-14)
    assert np.allclose(osp.diff_matrix(osp.cheb_points(3)), D3, atol=1e-14)
    assert np.allclose(osp.diff_matrix(x4), D4_ref, atol=1e-13)


@pytest.mark.parametrize("n", [5, 8, 16])
def test_diff_matrix_exact_on_polynomials(n):
    x = osp.cheb_points(n)
    D = osp.diff_matrix(x)
    for deg in range(n):
        assert np.allclose(D @ x**deg, deg * x ** max(deg - 1, 0) * (deg > 0), atol=1e-9 * n**2)


def test_lagrange_interpolation_examples():
    # interpolating from [-1, 1] to 0 gives [0.5, 0.5]
    assert np.allclose(osp.lagrange_matrix(np.array([-1.0, 1.0]), np.array([0.0])), [[0.5, 0.5]])
    x = osp.cheb_points(10)
    y = np.linspace(-1, 1, 7)
    M = osp.lagrange_matrix(x, y)
    for deg in range(10):
        assert np.allclose(M @ x**deg, y**deg, atol=1e-12)


def test_leaf_order_is_boundary_walk():
    p = 6
    perm, X, Y = osp.leaf_order_2D(p)
    nb = 4 * (p - 1)
    bx, by = X[:nb], Y[:nb]
    assert np.all((np.abs(bx) == 1) | (np.abs(by) == 1))
    assert np.all((np.abs(X[nb:]) < 1) & (np.abs(Y[nb:]) < 1))
    assert (bx[0], by[0]) == (-1.0, -1.0)  # starts at SW
    # consecutive boundary points are neighbours (counter-clockwise walk)
    steps = np.abs(np.diff(bx)) + np.abs(np.diff(by))
    assert np.all(steps > 0)


@pytest.mark.parametrize("eta", [None, 1.0])
def test_leaf_operators_exact(eta):
    """Q (normal derivative), G and QH (impedances) applied to a polynomial."""
    p, q = 8, 6
    ops = osp.leaf_ops_2D(p, q, 1.0, eta=eta)
    _, X, Y = osp.leaf_order_2D(p)
    u = X**3 + X * Y**2 - 2 * Y**3
    ux = 3 * X**2 + Y**2
    uy = 2 * X * Y - 6 * Y**2
    g = osp.gauss_nodes(q)
    gx = np.concatenate([g, np.ones(q), g[::-1], -np.ones(q)])
    gy = np.concatenate([-np.ones(q), g, np.ones(q), g[::-1]])
    un = np.concatenate(
        [-(2 * gx[:q] * gy[:q] - 6 * gy[:q] ** 2), 3 * gx[q:2 * q] ** 2 + gy[q:2 * q] ** 2,
         2 * gx[2 * q:3 * q] * gy[2 * q:3 * q] - 6 * gy[2 * q:3 * q] ** 2, -(3 * gx[3 * q:] ** 2 + gy[3 * q:] ** 2)]
    )
    assert np.allclose(ops["Q"] @ u, un, atol=1e-10)
    assert np.allclose(ops["D_x"] @ u, ux, atol=1e-10)
    assert np.allclose(ops["D_y"] @ u, uy, atol=1e-10)
    if eta is not None:
        ug = gx**3 + gx * gy**2 - 2 * gy**3
        assert np.allclose(ops["QH"] @ u, un - 1j * eta * ug, atol=1e-10)
        nb = 4 * (p - 1)
        # G: incoming impedance on the Chebyshev boundary (corners once)
        Xb, Yb = X[:nb], Y[:nb]
        assert ops["G"].shape == (nb, p * p)
        Gu = ops["G"] @ u
        assert np.allclose(Gu.imag, eta * u[:nb], atol=1e-10)


# ---------------------------------------------------------------------------
# closed-form end-to-end (the reference's accuracy cases)
# ---------------------------------------------------------------------------
def _setup(p, q, L, eta=None, bounds=cs.BOUNDS):
    pts = osp.leaf_points_2D(bounds, L, p)
    half = (bounds[1] - bounds[0]) / 2 ** (L + 1)
    ops = osp.leaf_ops_2D(p, q, half, eta=eta)
    rp = osp.boundary_gauss_points_2D(bounds, L, q)
    return pts, ops, rp


@pytest.mark.parametrize("L", [0, 1, 2])
def test_dtn_polynomial_closed_form(L):
    p, q = (16, 14) if L == 0 else (6, 4)
    pts, ops, rp = _setup(p, q, max(L, 0))
    ones = np.ones(pts.shape[:2])
    A = hps2d.assemble_diff_operator({"D_xx": ones, "D_yy": ones}, ops)
    f = cs.poly_source(pts)
    Y, T, v, h = hps2d.leaf_dtn(A, f[..., None], ops["P"], ops["Q"])
    g = cs.poly_dirichlet(rp)
    expected = cs.poly_dirichlet(pts) + cs.poly_part_soln(pts)
    if L == 0:
        u = Y[0] @ g + v[0, :, 0]
        assert np.abs(u - expected[0]).max() < 1e-12
        # DtN of the harmonic part gives its normal derivative; h = flux of v
        dn = cs.normal_derivative(rp, cs.poly_dirichlet_dx, cs.poly_dirichlet_dy)
        assert np.abs(T[0] @ g - dn).max() < 1e-11
        return
    S, gl, Ttop, _ = hps2d.merge_stage(T, h, L, "dtn")
    u, _ = hps2d.down_pass(g[:, None], S, gl, Y, v, "dtn")
    assert np.abs(u[..., 0] - expected).max() < 1e-12
    dn = cs.normal_derivative(rp, cs.poly_dirichlet_dx, cs.poly_dirichlet_dy)
    assert np.abs(Ttop @ g - dn).max() < 1e-11


@pytest.mark.parametrize("L", [1, 2])
def test_iti_complex_coefficient_closed_form(L):
    p, q = 6, 4
    pts, ops, rp = _setup(p, q, L, eta=cs.ETA)
    ones = np.ones(pts.shape[:2])
    A = hps2d.assemble_diff_operator({"D_xx": ones, "D_yy": ones, "I": cs.CPLX * ones}, ops)
    f = cs.cc_source(pts)
    T, Y, h, v = hps2d.leaf_iti(A, f[..., None], ops["P_iti"], ops["QH"], ops["G"])
    S, gl, Ttop, htop = hps2d.merge_stage(T, h, L, "iti")
    dn = cs.normal_derivative(rp, cs.cc_dx, cs.cc_dy)
    t_in = dn + 1j * cs.ETA * cs.cc_u(rp)
    u, _ = hps2d.down_pass(t_in[:, None], S, gl, Y, v, "iti")
    assert np.abs(u[..., 0] - cs.cc_u(pts)).max() < 1e-12
    # outgoing impedance of the solution = T t_in + h
    assert np.abs(Ttop @ t_in + htop[:, 0] - (dn - 1j * cs.ETA * cs.cc_u(rp))).max() < 1e-11


def test_iti_helmholtz_like_closed_form():
    """cases.py:149-209 at p = 16, q = 14, L = 1, atol 1e-8 (test_single_merge_accuracy.py:297-377)."""
    p, q, L = 16, 14, 1
    pts, ops, rp = _setup(p, q, L, eta=cs.ETA)
    ones = np.ones(pts.shape[:2])
    A = hps2d.assemble_diff_operator({"D_xx": ones, "D_yy": ones, "I": cs.hz_q(pts)}, ops)
    T, Y, h, v = hps2d.leaf_iti(A, cs.hz_f(pts)[..., None], ops["P_iti"], ops["QH"], ops["G"])
    S, gl, _, _ = hps2d.merge_stage(T, h, L, "iti")
    dn = cs.normal_derivative(rp, cs.hz_dx, cs.hz_dy)
    t_in = dn + 1j * cs.ETA * cs.hz_u(rp)
    u, _ = hps2d.down_pass(t_in[:, None], S, gl, Y, v, "iti")
    assert np.abs(u[..., 0] - cs.hz_u(pts)).max() < 1e-8


@pytest.mark.parametrize("kind", ["dtn", "iti"])
@pytest.mark.parametrize("L", [1, 2])
def test_stage_solution_equals_global_system(kind, L):
    p, q = 8, 6
    eta = 2.0 if kind == "iti" else None
    pts, ops, rp = _setup(p, q, L, eta=eta, bounds=(-1.0, 1.0, -1.0, 1.0))
    rng = np.random.default_rng(L)
    a = 1.0 + 0.3 * np.sin(2 * pts[..., 0]) * np.cos(pts[..., 1])
    c = 5.0 + np.cos(pts[..., 0] * pts[..., 1])
    coeffs = {"D_xx": a, "D_yy": a, "I": c}
    A = hps2d.assemble_diff_operator(coeffs, ops)
    f = np.sin(3 * pts[..., 0]) + pts[..., 1] ** 2
    g = rng.standard_normal(rp.shape[0])
    if kind == "dtn":
        Y, T, v, h = hps2d.leaf_dtn(A, f[..., None], ops["P"], ops["Q"])
        S, gl, _, _ = hps2d.merge_stage(T, h, L, "dtn")
        u, _ = hps2d.down_pass(g[:, None], S, gl, Y, v, "dtn")
        ug = gs.solve_dtn_global(A, f, ops["P"], ops["Q"], (-1.0, 1.0, -1.0, 1.0), L, q, g, rp)
    else:
        g = g + 1j * rng.standard_normal(rp.shape[0])
        T, Y, h, v = hps2d.leaf_iti(A, f[..., None].astype(complex), ops["P_iti"], ops["QH"], ops["G"])
        S, gl, _, _ = hps2d.merge_stage(T, h, L, "iti")
        u, _ = hps2d.down_pass(g[:, None], S, gl, Y, v, "iti")
        ug = gs.solve_iti_global(A, f, ops["P_iti"], ops["QH"], ops["G"], (-1.0, 1.0, -1.0, 1.0), L, q, g, rp)
    assert np.abs(u[..., 0] - ug).max() / np.abs(ug).max() < 1e-11


def test_nosource_iti_equals_with_source():
    p, q, L = 8, 6, 2
    pts, ops, rp = _setup(p, q, L, eta=1.0)
    ones = np.ones(pts.shape[:2])
    A = hps2d.assemble_diff_operator({"D_xx": ones, "D_yy": ones, "I": cs.CPLX * ones}, ops)
    f = cs.cc_source(pts)[..., None]
    T, Y, h, v = hps2d.leaf_iti(A, f, ops["P_iti"], ops["QH"], ops["G"])
    S, gl, _, _ = hps2d.merge_stage(T, h, L, "iti")
    Tn, Yn, Phi = hps2d.leaf_iti_nosource(A, ops["P_iti"], ops["QH"], ops["G"])
    Sn, Di, BDi, _ = hps2d.merge_stage_iti_nosource(Tn, L)
    nb = ops["P_iti"].shape[0]
    vv, gn = hps2d.up_pass_iti(f, Phi, ops["QH"], Di, BDi, nb)
    assert np.abs(vv - v).max() < 1e-12
    for a, b in zip(Sn, S):
        assert np.abs(a - b).max() < 1e-12
    for a, b in zip(gn, gl):
        assert np.abs(a - b).max() < 1e-10
