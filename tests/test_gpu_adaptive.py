"""GPU parity of the adaptive 2D DtN build + solve against the oracle, the
uniform path, and exact polynomial solutions."""

import numpy as np
import pytest
import torch

from oracle import adaptive2d as ad

pytestmark = pytest.mark.gpu


def _np(t):
    return t.detach().cpu().numpy()


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def _trees(kind, q):
    import jaxhps_amd as jh
    from jaxhps_amd import _adaptive as pa

    if kind == "ref":
        root = jh.DiscretizationNode2D(0.0, 1.0, 0.0, 1.0)
        pa.add_four_children(root, root=root, q=q)
        pa.add_four_children(root.children[0], root=root, q=q)
        pa.add_four_children(root.children[1], root=root, q=q)
        pa.add_four_children(root.children[0].children[1], root=root, q=q)
        o = ad.Node(0.0, 1.0, 0.0, 1.0)
        ad.split(o)
        ad.split(o.children[0])
        ad.split(o.children[1])
        ad.split(o.children[0].children[1])
        return root, o
    f = lambda x: np.exp(-30 * ((x[..., 0] - 0.15) ** 2 + (x[..., 1] - 0.1) ** 2))
    root = jh.DiscretizationNode2D(-1.0, 1.0, -1.0, 1.0)
    pa.generate_adaptive_mesh_level_restriction_2D(root, f, 1e-3, 6, q)
    o = ad.Node(-1.0, 1.0, -1.0, 1.0)
    ad.refine(o, f, 1e-3, 6)
    return root, o


@pytest.mark.parametrize("kind,p,q,nsrc", [("ref", 6, 4, None), ("deep", 8, 6, 3), ("deep", 12, 10, None)])
def test_adaptive_build_solve_matches_oracle(kind, p, q, nsrc):
    import jaxhps_amd as jh

    print(f"adaptive {kind} p={p} q={q} nsrc={nsrc}", flush=True)
    root, o = _trees(kind, q)
    dom = jh.Domain(p=p, q=q, root=root)
    rng = np.random.default_rng(p)
    shp = dom.interior_points.shape[:2]
    coeffs = {"D_xx": 1 + 0.3 * rng.random(shp), "D_yy": np.ones(shp), "D_x": 0.2 * rng.standard_normal(shp),
              "I": rng.standard_normal(shp)}
    src = rng.standard_normal(shp + ((nsrc,) if nsrc else ()))
    pde = jh.PDEProblem(dom, source=src, D_xx_coefficients=coeffs["D_xx"], D_yy_coefficients=coeffs["D_yy"],
                        D_x_coefficients=coeffs["D_x"], I_coefficients=coeffs["I"])
    T = jh.build_solver(pde, return_top_T=True)
    nb = dom.boundary_points.shape[0]
    assert tuple(T.shape) == (nb, nb)
    bdry = rng.standard_normal((nb,) + ((nsrc,) if nsrc else ()))
    u = _np(jh.solve(pde, torch.as_tensor(bdry)))
    uo, To = ad.solve(o, p, q, coeffs, src, bdry)
    assert _rel(_np(T), To) < 1e-9
    assert u.shape == shp + ((nsrc,) if nsrc else ())
    assert _rel(u if nsrc else u[..., None], uo) < 1e-9


def test_adaptive_path_on_uniform_tree_equals_uniform_path():
    import jaxhps_amd as jh
    from jaxhps_amd import _adaptive as pa

    p, q, L = 10, 8, 3
    root = jh.DiscretizationNode2D(0.0, 1.0, 0.0, 1.0)
    todo = [root]
    for _ in range(L):
        nxt = []
        for nd in todo:
            pa.add_four_children(nd, root=root, q=q)
            nxt += list(nd.children)
        todo = nxt
    dom_a = jh.Domain(p=p, q=q, root=root)
    dom_u = jh.Domain(p=p, q=q, root=jh.DiscretizationNode2D(0.0, 1.0, 0.0, 1.0), L=L)
    assert np.abs(dom_a.interior_points - dom_u.interior_points).max() < 1e-15
    rng = np.random.default_rng(1)
    shp = dom_u.interior_points.shape[:2]
    a, f = 1 + 0.5 * rng.random(shp), rng.standard_normal(shp)
    outs = []
    for dom in (dom_a, dom_u):
        pde = jh.PDEProblem(dom, source=f, D_xx_coefficients=a, D_yy_coefficients=np.ones(shp))
        T = _np(jh.build_solver(pde, return_top_T=True))
        g = np.cos(3 * dom.boundary_points[:, 0]) + dom.boundary_points[:, 1]
        outs.append((T, _np(jh.solve(pde, torch.as_tensor(g)))))
    assert _rel(outs[0][0], outs[1][0]) < 1e-10
    assert _rel(outs[0][1], outs[1][1]) < 1e-10


def test_polynomial_solution_exact_on_refined_tree():
    """u = x^3 - 2 x y^2 + y^3 + x is reproduced to rounding: leaf solves,
    compression and interface refinement are all exact at this degree."""
    import jaxhps_amd as jh

    p, q = 8, 6
    root, _ = _trees("deep", q)
    dom = jh.Domain(p=p, q=q, root=root)
    X, Y = dom.interior_points[..., 0], dom.interior_points[..., 1]
    u = X**3 - 2 * X * Y**2 + Y**3 + X
    a, c = 1 + X**2, np.sin(Y)
    f = a * 6 * X + (-4 * X + 6 * Y) + c * u
    pde = jh.PDEProblem(dom, source=f, D_xx_coefficients=a, D_yy_coefficients=np.ones_like(X), I_coefficients=c)
    jh.build_solver(pde)
    g_lst = dom.get_adaptive_boundary_data_lst(lambda x: x[:, 0] ** 3 - 2 * x[:, 0] * x[:, 1] ** 2 + x[:, 1] ** 3
                                               + x[:, 0])
    sol = _np(jh.solve(pde, g_lst))
    assert np.abs(sol - u).max() < 1e-10


def _bump_error(p, q, tol):
    import jaxhps_amd as jh

    uf = lambda x: np.exp(-25 * ((x[..., 0] - 0.2) ** 2 + (x[..., 1] + 0.1) ** 2))
    dom = jh.Domain.from_adaptive_discretization(p, q, jh.DiscretizationNode2D(-1.0, 1.0, -1.0, 1.0), uf, tol)
    X, Y = dom.interior_points[..., 0], dom.interior_points[..., 1]
    r2 = (X - 0.2) ** 2 + (Y + 0.1) ** 2
    lap = (2500 * r2 - 100) * np.exp(-25 * r2)
    pde = jh.PDEProblem(dom, source=lap, D_xx_coefficients=np.ones_like(X), D_yy_coefficients=np.ones_like(X))
    jh.build_solver(pde)
    sol = _np(jh.solve(pde, dom.get_adaptive_boundary_data_lst(uf)))
    return dom.n_leaves, np.abs(sol - uf(dom.interior_points)).max()


def test_from_adaptive_discretization_converges():
    """Poisson with a localised bump solution on refined trees: the error
    falls as the refinement tolerance does (FP64 oracle errors for these
    trees: 2.4e-3 at p = 10, tol 1e-4; 1.5e-6 at tol 1e-8; 2.7e-7 at
    p = 16, tol 1e-6)."""
    n1, e1 = _bump_error(10, 8, 1e-4)
    n2, e2 = _bump_error(10, 8, 1e-8)
    n3, e3 = _bump_error(16, 14, 1e-6)
    assert n2 > n1 and e2 < e1 / 100
    assert e1 < 5e-3 and e2 < 3e-6 and e3 < 1e-6
