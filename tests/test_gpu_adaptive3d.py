"""GPU parity of the adaptive 3D DtN build + solve against the oracle, the
uniform 3D path, and exact polynomial solutions."""

import numpy as np
import pytest
import torch

from oracle import adaptive3d as ad

pytestmark = pytest.mark.gpu

B0 = (0.0, 1.0, 0.0, 1.0, 0.0, 1.0)


def _np(t):
    return t.detach().cpu().numpy()


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def _trees(q):
    import jaxhps_amd as jh
    from jaxhps_amd import _adaptive3d as pa

    root = jh.DiscretizationNode3D(*B0)
    pa.add_eight_children(root, root=root, q=q)
    pa.add_eight_children(root.children[0], root=root, q=q)
    pa.add_eight_children(root.children[6], root=root, q=q)
    pa.add_eight_children(root.children[0].children[0], root=root, q=q)
    o = ad.Node(B0)
    ad.split(o)
    ad.split(o.children[0])
    ad.split(o.children[6])
    ad.split(o.children[0].children[0])
    return root, o


@pytest.mark.parametrize("p,q,nsrc", [(4, 2, None), (6, 4, 2)])
def test_adaptive3d_matches_oracle(p, q, nsrc):
    import jaxhps_amd as jh

    print(f"adaptive3d p={p} q={q} nsrc={nsrc}", flush=True)
    root, o = _trees(q)
    dom = jh.Domain(p=p, q=q, root=root)
    rng = np.random.default_rng(p)
    shp = dom.interior_points.shape[:2]
    coeffs = {"D_xx": 1 + 0.3 * rng.random(shp), "D_yy": np.ones(shp), "D_zz": np.ones(shp),
              "D_x": 0.2 * rng.standard_normal(shp), "I": rng.standard_normal(shp)}
    src = rng.standard_normal(shp + ((nsrc,) if nsrc else ()))
    pde = jh.PDEProblem(dom, source=src, D_xx_coefficients=coeffs["D_xx"], D_yy_coefficients=coeffs["D_yy"],
                        D_zz_coefficients=coeffs["D_zz"], D_x_coefficients=coeffs["D_x"],
                        I_coefficients=coeffs["I"])
    T = jh.build_solver(pde, return_top_T=True)
    nb = dom.boundary_points.shape[0]
    assert tuple(T.shape) == (nb, nb)
    bdry = rng.standard_normal((nb,) + ((nsrc,) if nsrc else ()))
    u = _np(jh.solve(pde, torch.as_tensor(bdry)))
    uo, To = ad.solve(o, p, q, coeffs, src, bdry)
    assert _rel(_np(T), To) < 1e-9
    assert _rel(u if nsrc else u[..., None], uo) < 1e-9


def test_adaptive3d_on_uniform_tree_equals_uniform_path():
    import jaxhps_amd as jh
    from jaxhps_amd import _adaptive3d as pa

    p, q, L = 6, 4, 2
    root = jh.DiscretizationNode3D(*B0)
    todo = [root]
    for _ in range(L):
        nxt = []
        for nd in todo:
            pa.add_eight_children(nd, root=root, q=q)
            nxt += list(nd.children)
        todo = nxt
    dom_a = jh.Domain(p=p, q=q, root=root)
    dom_u = jh.Domain(p=p, q=q, root=jh.DiscretizationNode3D(*B0), L=L)
    assert np.abs(dom_a.interior_points - dom_u.interior_points).max() < 1e-15
    assert np.abs(dom_a.boundary_points - dom_u.boundary_points).max() < 1e-15
    rng = np.random.default_rng(1)
    shp = dom_u.interior_points.shape[:2]
    a, f = 1 + 0.5 * rng.random(shp), rng.standard_normal(shp)
    outs = []
    for dom in (dom_a, dom_u):
        pde = jh.PDEProblem(dom, source=f, D_xx_coefficients=a, D_yy_coefficients=np.ones(shp),
                            D_zz_coefficients=np.ones(shp))
        T = _np(jh.build_solver(pde, return_top_T=True))
        g = np.cos(3 * dom.boundary_points[:, 0]) + dom.boundary_points[:, 1] * dom.boundary_points[:, 2]
        outs.append((T, _np(jh.solve(pde, torch.as_tensor(g)))))
    assert _rel(outs[0][0], outs[1][0]) < 1e-10
    assert _rel(outs[0][1], outs[1][1]) < 1e-10


def test_adaptive3d_polynomial_exact_and_refined_domain():
    import jaxhps_amd as jh

    p, q = 6, 4
    uf = lambda x: x[..., 0] ** 3 - 2 * x[..., 0] * x[..., 1] ** 2 + x[..., 2] ** 3 + x[..., 0] * x[..., 1] * x[..., 2]
    bump = lambda x: np.exp(-40 * ((x[..., 0] - 0.3) ** 2 + (x[..., 1] - 0.6) ** 2 + (x[..., 2] - 0.4) ** 2))
    dom = jh.Domain.from_adaptive_discretization(p, q, jh.DiscretizationNode3D(*B0), bump, 1e-2)
    assert dom.n_leaves > 8 and not dom.bool_uniform
    X, Y, Z = dom.interior_points[..., 0], dom.interior_points[..., 1], dom.interior_points[..., 2]
    a = 1 + X**2
    f = a * 6 * X + (-4 * X) + 6 * Z
    pde = jh.PDEProblem(dom, source=f, D_xx_coefficients=a, D_yy_coefficients=np.ones_like(X),
                        D_zz_coefficients=np.ones_like(X))
    jh.build_solver(pde)
    sol = _np(jh.solve(pde, dom.get_adaptive_boundary_data_lst(uf)))
    assert np.abs(sol - uf(dom.interior_points)).max() < 1e-9
