"""Pins of the 3D oracle: textbook/geometry facts, manufactured polynomial
solutions the discretisation reproduces exactly, and the dense global system
assembled from coordinates (no index tables)."""

import numpy as np
import pytest

from oracle import hps2d, hps3d

ORDER3 = ("D_xx", "D_xy", "D_yy", "D_xz", "D_yz", "D_zz", "D_x", "D_y", "D_z", "I")
CUBE = (-1.0, 1.0, -1.0, 1.0, -1.0, 1.0)


def _u(x):
    """Cubic with -Lap u = -(6x + 2z): reproduced exactly for q >= 4."""
    X, Y, Z = x[..., 0], x[..., 1], x[..., 2]
    return X**3 + Y**2 * Z - 2 * X * Y * Z + Z


def _f(x):
    return -(6 * x[..., 0] + 2 * x[..., 2])


def _setup(p, q, L, bounds=CUBE):
    half = (bounds[1] - bounds[0]) / 2 ** (L + 1)
    ops = hps3d.leaf_ops_3D(p, q, half)
    pts = hps3d.leaf_points_3D(bounds, L, p)
    rp = hps3d.boundary_gauss_points_3D(bounds, L, q)
    return ops, pts, rp


def test_leaf_operators_exact_on_polynomials():
    p, q = 6, 4
    ops, pts, _ = _setup(p, q, 0)
    x = pts[0]
    u = _u(x)
    X, Y, Z = x[..., 0], x[..., 1], x[..., 2]
    assert np.allclose(ops["D_x"] @ u, 3 * X**2 - 2 * Y * Z, atol=1e-12)
    assert np.allclose(ops["D_yz"] @ u, 2 * Y - 2 * X, atol=1e-11)
    lap = ops["D_xx"] @ u + ops["D_yy"] @ u + ops["D_zz"] @ u
    assert np.allclose(-lap, _f(x), atol=1e-10)
    # P reproduces a polynomial of degree < q from its Gauss face values
    gp = hps3d.leaf_gauss_points_3D(CUBE, q)
    nb = p**3 - (p - 2) ** 3
    assert np.abs(ops["P"] @ _u(gp) - u[:nb]).max() < 1e-12
    # Q gives the outward normal derivative at the Gauss points
    gx = np.stack([3 * gp[:, 0] ** 2 - 2 * gp[:, 1] * gp[:, 2],
                   2 * gp[:, 1] * gp[:, 2] - 2 * gp[:, 0] * gp[:, 2],
                   gp[:, 1] ** 2 - 2 * gp[:, 0] * gp[:, 1] + 1], -1)
    sgn = np.repeat([-1, 1, -1, 1, -1, 1], q * q)
    axis = np.repeat([0, 0, 1, 1, 2, 2], q * q)
    dn = sgn * gx[np.arange(gx.shape[0]), axis]
    assert np.abs(ops["Q"] @ u - dn).max() < 1e-10


@pytest.mark.parametrize("L", [1, 2])
def test_manufactured_cubic_exact(L):
    p, q = 6, 4
    ops, pts, rp = _setup(p, q, L)
    ones = np.ones(pts.shape[:2])
    A = hps2d.assemble_diff_operator({"D_xx": -ones, "D_yy": -ones, "D_zz": -ones}, ops, order=ORDER3)
    Y, T, v, h = hps2d.leaf_dtn(A, _f(pts)[..., None], ops["P"], ops["Q"])
    S, gl, Ttop, htop = hps3d.merge_stage_3d(T, h, L)
    u, _ = hps3d.down_pass_3d(_u(rp)[:, None], S, gl, Y, v)
    assert np.abs(u[..., 0] - _u(pts)).max() < 1e-11


def test_stage_solution_equals_global_system():
    p, q, L = 5, 3, 1
    ops, pts, rp = _setup(p, q, L)
    a = 1.0 + 0.3 * np.sin(2 * pts[..., 0]) * np.cos(pts[..., 1] + pts[..., 2])
    A = hps2d.assemble_diff_operator({"D_xx": -a, "D_yy": -a, "D_zz": -a, "I": 2.0 + 0 * a}, ops, order=ORDER3)
    f = np.sin(3 * pts[..., 0]) + pts[..., 1] * pts[..., 2]
    g = np.random.default_rng(1).standard_normal(rp.shape[0])
    Y, T, v, h = hps2d.leaf_dtn(A, f[..., None], ops["P"], ops["Q"])
    S, gl, _, _ = hps3d.merge_stage_3d(T, h, L)
    u, _ = hps3d.down_pass_3d(g[:, None], S, gl, Y, v)
    ug = hps3d.solve_dtn_global_3d(A, f, ops["P"], ops["Q"], CUBE, L, q, g, rp)
    assert np.abs(u[..., 0] - ug).max() / np.abs(ug).max() < 1e-11


def test_product_3d_operators_match_oracle():
    from jaxhps_amd import _precompute_operators as po

    for p, q in [(4, 2), (6, 4)]:
        half = 0.3
        ops = hps3d.leaf_ops_3D(p, q, half)
        d = po.precompute_diff_operators_3D(p, half)
        for a, nm in zip(d, ["D_x", "D_y", "D_z", "D_xx", "D_yy", "D_zz", "D_xy", "D_xz", "D_yz"]):
            assert np.allclose(a, ops[nm], atol=1e-10 * np.abs(ops[nm]).max()), nm
        assert np.allclose(po.precompute_P_3D_DtN(p, q), ops["P"], atol=1e-12)
        assert np.allclose(po.precompute_Q_3D_DtN(p, q, d[0], d[1], d[2]), ops["Q"], atol=1e-9 * np.abs(ops["Q"]).max())


def test_product_3d_points_match_oracle():
    import jaxhps_amd as jh

    for L in (0, 1, 2):
        dom = jh.Domain(p=4, q=2, root=jh.DiscretizationNode3D(*CUBE), L=L)
        assert np.allclose(dom.interior_points, hps3d.leaf_points_3D(CUBE, L, 4))
        assert np.allclose(dom.boundary_points, hps3d.boundary_gauss_points_3D(CUBE, L, 2))


@pytest.mark.parametrize("n", [1, 3])
def test_oct_tables_emulation_matches_oracle(n):
    from jaxhps_amd._tables import oct_merge_tables
    from test_host import _emulate_merge

    rng = np.random.default_rng(n)
    Tc = rng.standard_normal((8, 6 * n, 6 * n))
    hc = rng.standard_normal((8, 6 * n, 2))
    tb = oct_merge_tables(n)
    S, T, h, gt = _emulate_merge(Tc, hc, tb)
    S0, T0, h0, g0 = hps3d.oct_merge_dtn(Tc, hc)
    for a, b in [(S, S0), (T, T0), (h, h0), (gt, g0)]:
        assert np.abs(a - b).max() < 1e-10 * max(1.0, np.abs(b).max())
    g = rng.standard_normal((24 * n, 2))
    gint = S @ g + gt
    out = np.zeros((8, 6 * n, 2))
    for slot, src in enumerate(tb.child_src):
        c, l = divmod(slot, 6 * n)
        out[c, l] = g[src] if src >= 0 else gint[-src - 1]
    assert np.abs(out - hps3d.propagate_down_oct(S0, g, g0)).max() < 1e-10
