"""GPU parity of the 2D DtN / ItI build+solve against the FP64 oracle and
against the reference's closed-form accuracy cases."""

import numpy as np
import pytest
import torch

import cases2d as cs
from oracle import hps2d
from oracle import spectral as osp

pytestmark = pytest.mark.gpu


def _np(t):
    return t.detach().cpu().numpy()


def _problem(p, q, L, kind, nsrc=None, seed=0):
    import jaxhps_amd as jh

    root = jh.DiscretizationNode2D(*cs.BOUNDS)
    dom = jh.Domain(p=p, q=q, root=root, L=L)
    pts = dom.interior_points
    ones = np.ones(pts.shape[:2])
    rng = np.random.default_rng(seed)
    if kind == "dtn":
        f = cs.poly_source(pts)
        if nsrc:
            f = np.stack([f * (1 + 0.1 * k) for k in range(nsrc)], -1)
        pde = jh.PDEProblem(dom, source=f, D_xx_coefficients=ones, D_yy_coefficients=ones)
    else:
        f = cs.cc_source(pts)
        if nsrc:
            f = np.stack([f * (1 + 0.1 * k) for k in range(nsrc)], -1)
        pde = jh.PDEProblem(
            dom, source=f, D_xx_coefficients=ones, D_yy_coefficients=ones, I_coefficients=cs.CPLX * ones,
            use_ItI=True, eta=cs.ETA,
        )
    return dom, pde


def _oracle(pde, kind):
    dom = pde.domain
    half = (dom.root.xmax - dom.root.xmin) / 2 ** (dom.L + 1)
    ops = osp.leaf_ops_2D(dom.p, dom.q, half, eta=pde.eta if kind == "iti" else None)
    coeffs = {"D_xx": pde.D_xx_coefficients, "D_yy": pde.D_yy_coefficients, "I": pde.I_coefficients}
    A = hps2d.assemble_diff_operator(coeffs, ops)
    f = pde.source if pde.source.ndim == 3 else pde.source[..., None]
    if kind == "dtn":
        Y, T, v, h = hps2d.leaf_dtn(A, f, ops["P"], ops["Q"])
    else:
        T, Y, h, v = hps2d.leaf_iti(A, f, ops["P_iti"], ops["QH"], ops["G"])
    S_lst, g_lst, Ttop, htop = hps2d.merge_stage(T, h, dom.L, kind)
    return dict(Y=Y, T=T, v=v, h=h, S=S_lst, g=g_lst, Ttop=Ttop, A=A, ops=ops)


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("kind", ["dtn", "iti"])
@pytest.mark.parametrize("p,L", [(6, 1), (8, 2), (16, 1)])
def test_leaf_stage_matches_oracle(gpu, kind, p, L):
    from jaxhps_amd.local_solve import local_solve_stage_uniform_2D_DtN, local_solve_stage_uniform_2D_ItI

    print(f"leaf kind={kind} p={p} L={L}", flush=True)
    dom, pde = _problem(p, p - 2, L, kind)
    ref = _oracle(pde, kind)
    fn = local_solve_stage_uniform_2D_DtN if kind == "dtn" else local_solve_stage_uniform_2D_ItI
    Y, T, v, h = fn(pde)
    assert _rel(_np(Y), ref["Y"]) < 1e-11
    assert _rel(_np(T), ref["T"]) < 1e-11
    assert _rel(_np(v), ref["v"][..., 0]) < 1e-11
    assert _rel(_np(h), ref["h"][..., 0]) < 1e-11


@pytest.mark.parametrize("kind", ["dtn", "iti"])
@pytest.mark.parametrize("p,L", [(6, 1), (6, 2), (8, 3)])
def test_merge_stage_matches_oracle(gpu, kind, p, L):
    from jaxhps_amd.local_solve import local_solve_stage_uniform_2D_DtN, local_solve_stage_uniform_2D_ItI
    from jaxhps_amd.merge import merge_stage_uniform_2D_DtN, merge_stage_uniform_2D_ItI

    print(f"merge kind={kind} p={p} L={L}", flush=True)
    dom, pde = _problem(p, p - 2, L, kind)
    ref = _oracle(pde, kind)
    if kind == "dtn":
        Y, T, v, h = local_solve_stage_uniform_2D_DtN(pde)
        S_lst, g_lst, Ttop = merge_stage_uniform_2D_DtN(T, h, L, return_T=True)
    else:
        Y, T, v, h = local_solve_stage_uniform_2D_ItI(pde)
        S_lst, g_lst, Ttop = merge_stage_uniform_2D_ItI(T, h, L, return_T=True)
    assert len(S_lst) == L
    for lvl in range(L):
        assert tuple(S_lst[lvl].shape) == ref["S"][lvl].shape
        assert _rel(_np(S_lst[lvl]), ref["S"][lvl]) < 1e-10
        assert _rel(_np(g_lst[lvl]), ref["g"][lvl][..., 0]) < 1e-10
    assert _rel(_np(Ttop), ref["Ttop"]) < 1e-10


@pytest.mark.parametrize("p,L", [(6, 1), (8, 2), (16, 3)])
def test_dtn_end_to_end_closed_form(gpu, p, L):
    import jaxhps_amd as jh

    print(f"dtn e2e p={p} L={L}", flush=True)
    dom, pde = _problem(p, p - 2, L, "dtn")
    Ttop = jh.build_solver(pde, return_top_T=True)
    g = cs.poly_dirichlet(dom.boundary_points)
    u = _np(jh.solve(pde, g))
    expected = cs.poly_dirichlet(dom.interior_points) + cs.poly_part_soln(dom.interior_points)
    assert np.abs(u - expected).max() < 1e-11
    # top DtN map applied to the harmonic data gives its normal derivative
    dn = cs.normal_derivative(dom.boundary_points, cs.poly_dirichlet_dx, cs.poly_dirichlet_dy)
    assert np.abs(_np(Ttop) @ g - dn).max() < 1e-10


@pytest.mark.parametrize("p,L", [(6, 1), (8, 2), (16, 3)])
def test_iti_end_to_end_closed_form(gpu, p, L):
    import jaxhps_amd as jh

    print(f"iti e2e p={p} L={L}", flush=True)
    dom, pde = _problem(p, p - 2, L, "iti")
    jh.build_solver(pde)
    bp = dom.boundary_points
    dn = cs.normal_derivative(bp, cs.cc_dx, cs.cc_dy)
    t_in = dn + 1j * cs.ETA * cs.cc_u(bp)
    u = _np(jh.solve(pde, t_in))
    assert np.abs(u - cs.cc_u(dom.interior_points)).max() < 1e-10


@pytest.mark.parametrize("kind", ["dtn", "iti"])
def test_multi_source_and_oracle_solution(gpu, kind):
    import jaxhps_amd as jh

    print(f"multi-source kind={kind}", flush=True)
    nsrc = 3
    dom, pde = _problem(8, 6, 2, kind, nsrc=nsrc)
    ref = _oracle(pde, kind)
    jh.build_solver(pde)
    rng = np.random.default_rng(3)
    nbd = dom.boundary_points.shape[0]
    g = rng.standard_normal((nbd, nsrc)) + (1j * rng.standard_normal((nbd, nsrc)) if kind == "iti" else 0)
    u = _np(jh.solve(pde, g))
    assert u.shape == (dom.n_leaves, 64, nsrc)
    u_ref, _ = hps2d.down_pass(g, ref["S"], ref["g"], ref["Y"], ref["v"], kind)
    assert _rel(u, u_ref) < 1e-10
    with pytest.raises(ValueError):
        jh.solve(pde, g[:, 0])
