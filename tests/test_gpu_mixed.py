"""Mixed-precision build (FP32 / complex64 factors, FP64 top levels) with FP64
iterative refinement, against the FP64 GPU path (itself pinned to the oracle
in test_gpu_2d) and against the oracle directly on small trees."""

import numpy as np
import pytest
import torch

import cases2d as cs
from oracle import global_system as gs
from oracle import hps2d
from oracle import spectral as osp

pytestmark = pytest.mark.gpu


def _np(t):
    return t.detach().cpu().numpy()


def _rel2(a, b):
    return np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel())


def _last_update(hist):
    """The last measured or estimated relative update of the refinement."""
    h = hist[-1]
    return min(h["rel_update"], h.get("next_update_estimate", np.inf))


def _pair(make, **kw):
    """Two PDEProblems from the same inputs (one per precision)."""
    import jaxhps_amd as jh

    dom, kwargs, g = make(**kw)
    return dom, jh.PDEProblem(dom, **kwargs), jh.PDEProblem(dom, **kwargs), g


@pytest.mark.parametrize("L,k,top", [(3, 20.0, 2), (4, 40.0, 2), (4, 40.0, 0)])
def test_mixed_iti_helmholtz_refines_to_fp64(gpu, L, k, top):
    import jaxhps_amd as jh
    from jaxhps_amd import workloads as wl

    print(f"mixed iti L={L} k={k} top={top}", flush=True)
    dom, pde64, pdemx, g = _pair(wl.config_C, L=L, p=16, k=k, seed=1)
    jh.build_solver(pde64)
    u64 = _np(jh.solve(pde64, g))
    jh.build_solver(pdemx, precision="mixed", fp64_top_levels=top)
    lvl_dt = [str(s.dtype) for s in pdemx.S_lst]
    assert lvl_dt.count("torch.complex128") == min(top, L)
    u = jh.solve(pdemx, g, refine_tol=1e-12)
    assert str(u.dtype) == "torch.complex128"
    hist = pdemx.refine_history
    assert hist and _last_update(hist) <= 1e-12, hist
    assert _rel2(_np(u), u64) < 1e-9, (hist, _rel2(_np(u), u64))


@pytest.mark.parametrize("nsrc,low_leaves", [(1, None), (3, None), (3, True)])
def test_mixed_dtn_variable_coefficient_refines_to_fp64(gpu, nsrc, low_leaves):
    """2D DtN mixed build (FP64 leaves by default, FP32 leaves on request)."""
    import jaxhps_amd as jh
    from jaxhps_amd import workloads as wl

    print(f"mixed dtn nsrc={nsrc} low_leaves={low_leaves}", flush=True)
    dom, pde64, pdemx, g = _pair(wl.config_B, L=3, p=16, nsrc=nsrc, seed=2)
    rng = np.random.default_rng(4)
    f = rng.standard_normal(pde64.source.shape)
    pde64.source = f
    pdemx.source = f
    jh.build_solver(pde64)
    u64 = _np(jh.solve(pde64, g))
    jh.build_solver(pdemx, precision="mixed", low_precision_leaves=low_leaves)
    leaf_dt = pdemx._factors["leaf"].lu.dtype
    assert leaf_dt == (torch.float32 if low_leaves else torch.float64), leaf_dt
    u = jh.solve(pdemx, g, refine_tol=1e-12)
    assert str(u.dtype) == "torch.float64"
    hist = pdemx.refine_history
    assert _last_update(hist) <= 1e-12, hist
    assert _rel2(_np(u), u64) < 1e-10, (hist, _rel2(_np(u), u64))
    # the same through solve(source=...) on the mixed factors
    u2 = jh.solve(pdemx, g, source=2 * f, refine_tol=1e-12)
    pde64.source = 2 * f
    jh.build_solver(pde64)
    assert _rel2(_np(u2), _np(jh.solve(pde64, g))) < 1e-10


def test_mixed_iti_matches_oracle_global_system(gpu):
    """Small complex-coefficient ItI case: mixed+refined equals the dense global
    solve of the oracle (and the reference's closed form to discretisation error)."""
    import jaxhps_amd as jh

    print("mixed iti vs global system", flush=True)
    L, p, q = 2, 8, 6
    dom = jh.Domain(p=p, q=q, root=jh.DiscretizationNode2D(*cs.BOUNDS), L=L)
    pts = dom.interior_points
    ones = np.ones(pts.shape[:2])
    f = cs.cc_source(pts)
    pde = jh.PDEProblem(dom, source=f, D_xx_coefficients=ones, D_yy_coefficients=ones,
                        I_coefficients=cs.CPLX * ones, use_ItI=True, eta=cs.ETA)
    jh.build_solver(pde, precision="mixed", fp64_top_levels=1)
    bp = dom.boundary_points
    t_in = cs.normal_derivative(bp, cs.cc_dx, cs.cc_dy) + 1j * cs.ETA * cs.cc_u(bp)
    u = _np(jh.solve(pde, t_in, refine_tol=1e-13))
    half = (dom.root.xmax - dom.root.xmin) / 2 ** (L + 1)
    ops = osp.leaf_ops_2D(p, q, half, eta=cs.ETA)
    A = hps2d.assemble_diff_operator({"D_xx": ones, "D_yy": ones, "I": cs.CPLX * ones}, ops)
    u_ref = gs.solve_iti_global(A, f, ops["P_iti"], ops["QH"], ops["G"], cs.BOUNDS, L, q, t_in, bp)
    assert _rel2(u, u_ref) < 1e-11
    assert np.abs(u - cs.cc_u(pts)).max() < 1e-9


def test_mixed_nosource_iti_with_source_at_solve(gpu):
    import jaxhps_amd as jh
    from jaxhps_amd import workloads as wl

    print("mixed nosource iti", flush=True)
    dom, kwargs, g = wl.config_C(L=3, p=12, k=20.0, seed=3)
    rng = np.random.default_rng(8)
    f = rng.standard_normal(dom.interior_points.shape[:2]) + 1j * rng.standard_normal(dom.interior_points.shape[:2])
    kw = dict(kwargs)
    kw.pop("source")
    pde = jh.PDEProblem(dom, **kw)
    jh.build_solver(pde, precision="mixed")
    u = _np(jh.solve(pde, g, source=f, refine_tol=1e-12))
    kw["source"] = f
    pde64 = jh.PDEProblem(dom, **kw)
    jh.build_solver(pde64)
    assert _rel2(u, _np(jh.solve(pde64, g))) < 1e-10


def test_fp64_solution_has_zero_residual(gpu):
    """The FP64 HPS solution satisfies the global system the residual kernels
    evaluate (pins their neighbour and row conventions): its residual is
    rounding-sized next to the residual change a 1e-6 perturbation of one
    leaf makes, and that change stays on the leaf and its edge neighbours."""
    import torch

    import jaxhps_amd as jh
    from jaxhps_amd import _refine as rf
    from jaxhps_amd import workloads as wl
    from jaxhps_amd._neighbors import gauss_neighbors_2D
    from jaxhps_amd.down_pass import _down

    print("residual of fp64 solution", flush=True)
    for make, kw in [(wl.config_C, dict(L=3, p=10, k=15.0, seed=5)), (wl.config_B, dict(L=3, p=10, nsrc=2, seed=5))]:
        dom, kwargs, g = make(**kw)
        pde = jh.PDEProblem(dom, **kwargs)
        jh.build_solver(pde)
        u, g_leaf = _down(g, pde.S_lst, pde.g_tilde_lst, pde.Y, pde.v, return_leaf_g=True)
        u = (u if u.ndim == 3 else u.unsqueeze(-1)).contiguous()
        dt = u.dtype
        ops = rf.refine_ops(pde, u.device, dt)
        f = torch.as_tensor(kwargs["source"]).to(u.device, dt)
        f = (f if f.ndim == 3 else f.unsqueeze(-1)).contiguous()
        groot = rf._root_data(g, u.device, dt, u.shape[-1])
        nb = pde.P.shape[0]
        du = torch.zeros_like(u)
        du[7, nb:, :] = 1e-6 * u.abs().max()
        up = (u + du).contiguous()

        def res(x):
            if pde.use_ItI:
                return [_np(rf.residual_iti(ops, x, f, groot, dt)[0])]
            r, hinj, gr, _ = rf.residual_dtn(ops, x, g_leaf.contiguous(), f, groot, dt)
            return [_np(r), _np(hinj), _np(gr)]

        r0, r1 = res(u), res(up)
        d = [b - a for a, b in zip(r0, r1)]
        big = max(np.abs(x).max() for x in d)
        assert max(np.abs(x).max() for x in r0) < 1e-4 * big
        touched = {7} | {int(s) // (4 * dom.q) for s in gauss_neighbors_2D(dom.L, dom.q)[7] if s >= 0}
        for l in range(dom.n_leaves):
            if l not in touched:
                assert np.abs(d[0][l]).max() == 0.0, l
        if not pde.use_ItI:
            assert np.abs(d[2]).max() == 0.0


def test_refinement_stop_estimate_matches_the_next_update(gpu):
    """The refinement stops on an estimate of the next update (the last update
    times the residual's reduction).  Running the same solve with a tighter
    tolerance computes that update: the estimate must be within a factor of
    ten of it (config C's workload, k = 100, L = 4, complex64 leaves and
    merges)."""
    import jaxhps_amd as jh
    from jaxhps_amd import workloads as wl

    print("refinement estimate vs actual update", flush=True)
    dom, kw, g = wl.config_C(L=4)
    pde = jh.PDEProblem(dom, **kw)
    jh.build_solver(pde, precision="mixed")
    jh.solve(pde, g, refine_tol=1e-10)
    h1 = list(pde.refine_history)
    assert pde.refine_converged and "next_update_estimate" in h1[-1]
    est, n1 = h1[-1]["next_update_estimate"], len(h1)
    jh.solve(pde, g, refine_tol=1e-15, max_refine=n1 + 1)
    h2 = pde.refine_history
    assert len(h2) > n1
    for a, b in zip(h1, h2):
        assert a["rel_update"] == b["rel_update"]  # the same iterations
    actual = h2[n1]["rel_update"]
    assert est / 10 <= actual <= est * 10, (est, actual)
