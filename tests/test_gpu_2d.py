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


def test_leaf_natural_order_rejected_falls_back_to_pivoting(gpu):
    """The leaf stage factors in natural order with a per-leaf multiplier
    check; an I coefficient that cancels the first interior diagonal entry of
    A_ii (a zero or tiny first pivot in natural order, while A_ii stays
    nonsingular) must send the stage back to partial pivoting, whose result
    matches the oracle's LAPACK inverses."""
    import jaxhps_amd as jh
    from jaxhps_amd.local_solve import local_solve_stage_uniform_2D_DtN

    p, q, L = 8, 6, 1
    root = jh.DiscretizationNode2D(*cs.BOUNDS)
    dom = jh.Domain(p=p, q=q, root=root, L=L)
    pts = dom.interior_points
    ones = np.ones(pts.shape[:2])
    half = (dom.root.xmax - dom.root.xmin) / 2 ** (L + 1)
    ops = osp.leaf_ops_2D(p, q, half)
    nb = ops["P"].shape[0]
    I = np.zeros(pts.shape[:2])
    I[:, nb] = -(ops["D_xx"][nb, nb] + ops["D_yy"][nb, nb])
    pde = jh.PDEProblem(dom, source=cs.poly_source(pts), D_xx_coefficients=ones, D_yy_coefficients=ones,
                        I_coefficients=I)
    print("leaf natural-order fallback", flush=True)
    ref = _oracle(pde, "dtn")
    Y, T, v, h = local_solve_stage_uniform_2D_DtN(pde)
    assert _rel(_np(Y), ref["Y"]) < 1e-9
    assert _rel(_np(T), ref["T"]) < 1e-9
    assert _rel(_np(v), ref["v"][..., 0]) < 1e-9


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


def test_iti_block_sparse_couplings_match_oracle(gpu):
    """Merges whose children have >= 64 nodes per side multiply only the
    nonzero rectangles of D12 / D21; fed the oracle's leaf maps, the GPU merge
    stage must reproduce the oracle's dense one to rounding."""
    from jaxhps_amd._tables import quad_merge_tables
    from jaxhps_amd.merge import merge_stage_uniform_2D_ItI

    print("iti block-sparse couplings p=16 L=4", flush=True)
    assert len(quad_merge_tables(112, "iti").d12_blk) > 0
    dom, pde = _problem(16, 14, 4, "iti")
    ref = _oracle(pde, "iti")
    T = torch.as_tensor(ref["T"]).to(gpu)
    h = torch.as_tensor(ref["h"][..., 0]).to(gpu)
    S_lst, g_lst, Ttop = merge_stage_uniform_2D_ItI(T, h, 4, return_T=True)
    for lvl in range(4):
        assert _rel(_np(S_lst[lvl]), ref["S"][lvl]) < 1e-11
        assert _rel(_np(g_lst[lvl]), ref["g"][lvl][..., 0]) < 1e-11
    assert _rel(_np(Ttop), ref["Ttop"]) < 1e-11


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
    # At p = 16, L = 3 rounding dominates the discretisation error; the leaf
    # matrices are factored with their impedance rows scaled to the size of
    # the collocation rows (condition ~3e5 -> ~3e4).
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


@pytest.mark.parametrize("kind", ["dtn", "iti"])
@pytest.mark.parametrize("nsrc", [None, 2])
def test_up_pass_new_source_matches_rebuild(gpu, kind, nsrc):
    """solve(pde, g, source=f2) on a solver built for f1 equals a fresh build for f2."""
    import jaxhps_amd as jh

    print(f"up pass kind={kind} nsrc={nsrc}", flush=True)
    dom, pde1 = _problem(8, 6, 3, kind, nsrc=nsrc, seed=0)
    _, pde2 = _problem(8, 6, 3, kind, nsrc=nsrc, seed=0)
    rng = np.random.default_rng(11)
    f2 = rng.standard_normal(pde1.source.shape)
    pde2.source = f2 if kind == "dtn" else f2 + 0.5j * rng.standard_normal(f2.shape)
    jh.build_solver(pde1)
    jh.build_solver(pde2)
    nb = dom.boundary_points.shape[0]
    g = rng.standard_normal((nb,) + ((nsrc,) if nsrc else ()))
    if kind == "iti":
        g = g + 1j * rng.standard_normal(g.shape)
    u_new = _np(jh.solve(pde1, g, source=pde2.source))
    u_ref = _np(jh.solve(pde2, g))
    assert _rel(u_new, u_ref) < 1e-11
    from jaxhps_amd.up_pass import up_pass_uniform_2D_DtN, up_pass_uniform_2D_ItI

    fn = up_pass_uniform_2D_ItI if kind == "iti" else up_pass_uniform_2D_DtN
    v, g_lst = fn(pde2.source, pde1)
    assert _rel(_np(v), _np(pde2.v)) < 1e-11
    for a, b in zip(g_lst, pde2.g_tilde_lst):
        assert _rel(_np(a), _np(b)) < 1e-10


@pytest.mark.parametrize("L", [1, 3])
def test_nosource_iti_build_then_source_solve(gpu, L):
    """No-source ItI build + solve(source=f) against the oracle's explicit-inverse
    up pass (Phi, D^{-1}, B D^{-1}) and against the with-source build."""
    import jaxhps_amd as jh

    print(f"nosource iti L={L}", flush=True)
    dom, pde_src = _problem(8, 6, L, "iti", nsrc=2)
    f = pde_src.source
    ones = np.ones(dom.interior_points.shape[:2])
    pde = jh.PDEProblem(dom, D_xx_coefficients=ones, D_yy_coefficients=ones, I_coefficients=cs.CPLX * ones,
                        use_ItI=True, eta=cs.ETA)
    Ttop = jh.build_solver(pde, return_top_T=True)
    jh.build_solver(pde_src)
    rng = np.random.default_rng(5)
    g = rng.standard_normal((dom.boundary_points.shape[0], 2)) + 1j * rng.standard_normal((dom.boundary_points.shape[0], 2))
    u = _np(jh.solve(pde, g, source=f))
    assert _rel(u, _np(jh.solve(pde_src, g))) < 1e-11
    # oracle: explicit inverses
    half = (dom.root.xmax - dom.root.xmin) / 2 ** (L + 1)
    ops = osp.leaf_ops_2D(dom.p, dom.q, half, eta=cs.ETA)
    coeffs = {"D_xx": ones, "D_yy": ones, "I": cs.CPLX * ones}
    A = hps2d.assemble_diff_operator(coeffs, ops)
    T0, Y0, Phi = hps2d.leaf_iti_nosource(A, ops["P_iti"], ops["QH"], ops["G"])
    S0, Di, BDi, Tt0 = hps2d.merge_stage_iti_nosource(T0, L)
    assert _rel(_np(Ttop), Tt0) < 1e-10
    for a, b in zip(pde.S_lst, S0):
        assert _rel(_np(a), b) < 1e-10
    v0, g0 = hps2d.up_pass_iti(f, Phi, ops["QH"], Di, BDi, ops["P_iti"].shape[0])
    from jaxhps_amd.up_pass import up_pass_uniform_2D_ItI

    v, g_lst = up_pass_uniform_2D_ItI(f, pde)
    assert _rel(_np(v), v0) < 1e-11
    for a, b in zip(g_lst, g0):
        assert _rel(_np(a), b) < 1e-10


@pytest.mark.parametrize("kind", ["dtn", "iti"])
def test_subtree_recomputation_matches_build_solve(gpu, kind):
    """solve_subtree and upward/downward_pass_subtree (_subtree_recomp.py)
    give the oracle's solution and the build_solver + solve one."""
    import jaxhps_amd as jh

    print(f"subtree recomputation kind={kind}", flush=True)
    dom, pde = _problem(8, 6, 4, kind, seed=2)
    rng = np.random.default_rng(4)
    nb = dom.boundary_points.shape[0]
    g = rng.standard_normal(nb) + (1j * rng.standard_normal(nb) if kind == "iti" else 0)
    # the oracle's stage-by-stage solution of the same problem
    ref = _oracle(pde, kind)
    u_oracle, _ = hps2d.down_pass(g[:, None], ref["S"], ref["g"], ref["Y"], ref["v"], kind)
    jh.build_solver(pde)
    u_ref = _np(jh.solve(pde, g))
    assert _rel(u_ref, u_oracle[..., 0]) < 1e-10
    _, pde2 = _problem(8, 6, 4, kind, seed=2)
    u_sub = _np(jh.solve_subtree(pde2, g, subtree_height=2))
    assert _rel(u_sub, u_oracle[..., 0]) < 1e-10
    assert _rel(u_sub, u_ref) < 1e-11
    _, pde3 = _problem(8, 6, 4, kind, seed=2)
    T = jh.upward_pass_subtree(pde3, subtree_height=2)
    assert T.shape == (nb, nb)
    u_ud = _np(jh.downward_pass_subtree(pde3, g, subtree_height=2))
    assert _rel(u_ud, u_oracle[..., 0]) < 1e-10
    assert _rel(u_ud, u_ref) < 1e-11
    # all-together branch when the tree is no taller than the subtree height
    _, pde4 = _problem(8, 6, 4, kind, seed=2)
    assert _rel(_np(jh.solve_subtree(pde4, g, subtree_height=4)), u_ref) < 1e-11


def test_nosource_dtn_build_then_source_solve(gpu):
    """No-source DtN build (_build_solver.py:261-331 with the DtN stages) then
    solve(source=f) equals the with-source build."""
    import jaxhps_amd as jh

    print("nosource dtn", flush=True)
    dom, pde_src = _problem(8, 6, 3, "dtn", nsrc=2)
    ones = np.ones(dom.interior_points.shape[:2])
    pde = jh.PDEProblem(dom, D_xx_coefficients=ones, D_yy_coefficients=ones)
    Ttop = jh.build_solver(pde, return_top_T=True)
    Ttop_src = jh.build_solver(pde_src, return_top_T=True)
    assert _rel(_np(Ttop), _np(Ttop_src)) < 1e-11
    g = np.random.default_rng(7).standard_normal((dom.boundary_points.shape[0], 2))
    u = _np(jh.solve(pde, g, source=pde_src.source))
    assert _rel(u, _np(jh.solve(pde_src, g))) < 1e-11


def test_unpivoted_merge_falls_back_to_pivoting(gpu):
    """A DtN merge whose interface matrix (1120 unknowns) has an exactly zero
    leading entry (nonsingular otherwise) breaks the unpivoted factorisation; the factor
    check must catch it and the level be redone with partial pivoting, giving
    the oracle's result.  The regular merge matrices of the workloads take
    the unpivoted path (every other parity test covers it)."""
    from jaxhps_amd._stages import merge_factors_stable
    from jaxhps_amd.merge import _levels, merge_stage_uniform_2D_DtN
    from jaxhps_amd._tables import quad_merge_tables

    print("unpivoted merge fallback", flush=True)
    m = 280  # 1120 interface unknowns: taller than the register panels, so unpivoted
    ne = 4 * m
    rng = np.random.default_rng(21)
    T = 2.0 * np.eye(ne) + rng.standard_normal((4, ne, ne)) / np.sqrt(ne)
    h = rng.standard_normal((4, ne, 1))
    # D[0, 0] = a55[0, 0] + b55[0, 0]: a's interface 5 is its E side (rows m..2m),
    # b's is its W side, reversed (rows 4m-1 .. 3m)
    T[1, ne - 1, ne - 1] = -T[0, m, m]
    S0, g0, Tt0, _ = hps2d.merge_stage(T, h, 1, "dtn")
    Tg = torch.as_tensor(T, device=gpu)
    hg = torch.as_tensor(h, device=gpu)
    # unpivoted factors of this level fail the check ...
    levels, _ = _levels(Tg, hg, 1, "dtn", 4, lambda nec: quad_merge_tables(nec // 4, "dtn"), pivot=True)
    assert bool(merge_factors_stable(levels[0].F_lu))
    from jaxhps_amd._stages import merge_level

    raw = merge_level(Tg.contiguous(), hg.contiguous(), quad_merge_tables(m, "dtn"), pivot=False)
    assert not bool(merge_factors_stable(raw.F_lu))
    # ... and the merge stage redoes it with pivoting
    S_lst, g_lst, Ttop = merge_stage_uniform_2D_DtN(Tg, hg, 1, return_T=True)
    assert _rel(_np(S_lst[0]), S0[0]) < 1e-11
    assert _rel(_np(Ttop), Tt0) < 1e-11


def test_unpivoted_register_panel_merge_rejected_falls_back(gpu):
    """32 DtN merges of 128 interface unknowns (register-sized panels, more
    than 16 matrices): merge 5's interface matrix has an exactly zero
    leading entry.  The natural-order panel kernel must flag that matrix
    alone (info = -1) and the merge stage redo the level with partial
    pivoting, giving the oracle's S for every merge."""
    from jaxhps_amd._stages import merge_level
    from jaxhps_amd._tables import quad_merge_tables
    from jaxhps_amd.merge import merge_stage_uniform_2D_DtN

    print("unpivoted register-panel merge fallback", flush=True)
    m = 32
    ne = 4 * m
    rng = np.random.default_rng(23)
    T = 2.0 * np.eye(ne) + rng.standard_normal((128, ne, ne)) / np.sqrt(ne)
    h = rng.standard_normal((128, ne, 1))
    bad = 5
    # D[0, 0] = a55[0, 0] + b55[0, 0] of merge `bad` (children 4 bad, 4 bad + 1)
    T[4 * bad + 1, ne - 1, ne - 1] = -T[4 * bad, m, m]
    S0, _, _, _ = hps2d.merge_stage(T, h, 1, "dtn")
    Tg = torch.as_tensor(T, device=gpu)
    hg = torch.as_tensor(h, device=gpu)
    raw = merge_level(Tg, hg, quad_merge_tables(m, "dtn"), pivot=False)
    info = _np(raw.info)
    assert info[bad] == -1 and (np.delete(info, bad) == 0).all(), info
    S_lst, _ = merge_stage_uniform_2D_DtN(Tg, hg, 2)
    assert _rel(_np(S_lst[0]), S0[0]) < 1e-11


@pytest.mark.parametrize("kind,m,nsrc", [("dtn", 80, 1), ("dtn", 80, 4), ("iti", 72, 1), ("iti", 72, 16)])
def test_up_pass_through_block_inverses_matches_blocked_solve(gpu, kind, m, nsrc):
    """A low-precision merge level of 3 matrices of 320 (DtN) / 288 (ItI)
    unknowns -- one full 256-block and a partial one -- keeps block inverses;
    the up pass through them gives the blocked triangular solve's g_tilde and
    h to single precision, and both match the complex128 / float64 up pass.
    The merge's own source solve (through inverses it forms in its
    workspace) matches the double-precision merge too."""
    from jaxhps_amd._stages import form_block_inverses, merge_level, up_level
    from jaxhps_amd._tables import quad_merge_tables

    print(f"up pass block inverses kind={kind} m={m} nsrc={nsrc}", flush=True)
    rng = np.random.default_rng(31)
    nm, nec = 3, 4 * m
    T = 2.0 * np.eye(nec) + rng.standard_normal((4 * nm, nec, nec)) / np.sqrt(nec)
    h = rng.standard_normal((4 * nm, nec, nsrc))
    if kind == "iti":
        T = T + 1j * rng.standard_normal(T.shape) / np.sqrt(nec)
        h = h + 1j * rng.standard_normal(h.shape)
    lo, hi = (torch.float32, torch.float64) if kind == "dtn" else (torch.complex64, torch.complex128)
    tb = quad_merge_tables(m, kind)
    res = {}
    for dt in (lo, hi):
        Tg = torch.as_tensor(T, device=gpu).to(dt).contiguous()
        hg = torch.as_tensor(h, device=gpu).to(dt).contiguous()
        lvl = merge_level(Tg, hg, tb, keep_factors=True)
        res[dt, "merge"] = (lvl.gt, lvl.h)
        assert (lvl.Finv is not None) == (dt == lo)
        if lvl.Finv is not None:  # the merge's inverses are hps_block_inverses'
            from_merge = lvl.Finv.clone()
            form_block_inverses(lvl)
            assert torch.equal(lvl.Finv, from_merge)
        res[dt, True] = up_level(hg, lvl)
        lvl.Finv = None
        res[dt, False] = up_level(hg, lvl)
    for i in range(2):  # g_tilde, h_out
        ref = _np(res[hi, "merge"][i])
        assert _rel(_np(res[lo, "merge"][i]), ref) < 1e-4
        ref = _np(res[hi, False][i])
        assert _rel(_np(res[hi, True][i]), ref) < 1e-14
        assert _rel(_np(res[lo, True][i]), _np(res[lo, False][i])) < 1e-4
        assert _rel(_np(res[lo, True][i]), ref) < 1e-4
