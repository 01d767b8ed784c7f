"""GPU parity of the 3D DtN build+solve (8:1 merges) against the FP64 oracle
and a manufactured solution the discretisation reproduces exactly."""

import numpy as np
import pytest

from oracle import hps2d, hps3d

pytestmark = pytest.mark.gpu

ORDER3 = ("D_xx", "D_xy", "D_yy", "D_xz", "D_yz", "D_zz", "D_x", "D_y", "D_z", "I")
CUBE = (-1.0, 1.0, -1.0, 1.0, -1.0, 1.0)


def _np(t):
    return t.detach().cpu().numpy()


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def _u(x):
    X, Y, Z = x[..., 0], x[..., 1], x[..., 2]
    return X**3 + Y**2 * Z - 2 * X * Y * Z + Z


def _f(x):
    return -(6 * x[..., 0] + 2 * x[..., 2])


def _varcoef_problem(p, L, nsrc=None, seed=0):
    import jaxhps_amd as jh

    dom = jh.Domain(p=p, q=p - 2, root=jh.DiscretizationNode3D(*CUBE), L=L)
    x = dom.interior_points
    a = 1.0 + 0.3 * np.sin(2 * x[..., 0]) * np.cos(x[..., 1] + x[..., 2])
    rng = np.random.default_rng(seed)
    f = np.sin(3 * x[..., 0]) + x[..., 1] * x[..., 2]
    if nsrc:
        f = np.stack([f * (1 + 0.2 * k) for k in range(nsrc)], -1)
    pde = jh.PDEProblem(dom, source=f, D_xx_coefficients=-a, D_yy_coefficients=-a, D_zz_coefficients=-a,
                        D_x_coefficients=0.1 * a, I_coefficients=2.0 + 0 * a)
    half = (CUBE[1] - CUBE[0]) / 2 ** (L + 1)
    ops = hps3d.leaf_ops_3D(p, p - 2, half)
    A = hps2d.assemble_diff_operator({"D_xx": -a, "D_yy": -a, "D_zz": -a, "D_x": 0.1 * a, "I": 2.0 + 0 * a}, ops,
                                     order=ORDER3)
    g = rng.standard_normal((dom.boundary_points.shape[0],) + ((nsrc,) if nsrc else ()))
    return dom, pde, ops, A, f, g


@pytest.mark.parametrize("p,L", [(5, 1), (6, 2)])
def test_3d_stages_match_oracle(gpu, p, L):
    from jaxhps_amd.local_solve import local_solve_stage_uniform_3D_DtN
    from jaxhps_amd.merge import merge_stage_uniform_3D_DtN

    print(f"3d stages p={p} L={L}", flush=True)
    dom, pde, ops, A, f, g = _varcoef_problem(p, L)
    Y0, T0, v0, h0 = hps2d.leaf_dtn(A, f[..., None], ops["P"], ops["Q"])
    S0, g0, Tt0, _ = hps3d.merge_stage_3d(T0, h0, L)
    Y, T, v, h = local_solve_stage_uniform_3D_DtN(pde)
    assert _rel(_np(Y), Y0) < 1e-11 and _rel(_np(T), T0) < 1e-11
    assert _rel(_np(v), v0[..., 0]) < 1e-11 and _rel(_np(h), h0[..., 0]) < 1e-11
    S_lst, g_lst, Ttop = merge_stage_uniform_3D_DtN(T, h, L, return_T=True)
    assert len(S_lst) == L
    assert S_lst[-1].ndim == 2  # the reference stores the root without a batch axis
    for lvl in range(L):
        ref = S0[lvl] if lvl < L - 1 else S0[lvl][0]
        assert tuple(S_lst[lvl].shape) == ref.shape
        assert _rel(_np(S_lst[lvl]), ref) < 1e-10
        gref = g0[lvl][..., 0] if lvl < L - 1 else g0[lvl][0, :, 0]
        assert _rel(_np(g_lst[lvl]), gref) < 1e-10
    assert _rel(_np(Ttop), Tt0) < 1e-10


@pytest.mark.parametrize("L", [1, 2])
def test_3d_manufactured_cubic_exact(gpu, L):
    import jaxhps_amd as jh

    print(f"3d cubic L={L}", flush=True)
    dom = jh.Domain(p=6, q=4, root=jh.DiscretizationNode3D(*CUBE), L=L)
    x = dom.interior_points
    ones = np.ones(x.shape[:2])
    pde = jh.PDEProblem(dom, source=_f(x), D_xx_coefficients=-ones, D_yy_coefficients=-ones,
                        D_zz_coefficients=-ones)
    jh.build_solver(pde)
    u = _np(jh.solve(pde, _u(dom.boundary_points)))
    assert np.abs(u - _u(x)).max() < 1e-11


def test_3d_solution_matches_global_system_and_multisource(gpu):
    import jaxhps_amd as jh

    print("3d global + multi-source", flush=True)
    p, L = 5, 1
    dom, pde, ops, A, f, g = _varcoef_problem(p, L, nsrc=2)
    jh.build_solver(pde)
    u = _np(jh.solve(pde, g))
    assert u.shape == (8, p**3, 2)
    for k in range(2):
        ug = hps3d.solve_dtn_global_3d(A, f[..., k], ops["P"], ops["Q"], CUBE, L, p - 2, g[:, k],
                                       dom.boundary_points)
        assert _rel(u[..., k], ug) < 1e-10
    with pytest.raises(ValueError):
        jh.solve(pde, g[:, 0])


def test_config_d_workload_small(gpu):
    """Config D's problem (bumps coefficient, manufactured u, 4 RHS) at L=1,
    p=8: matches the oracle stage-by-stage solution."""
    import jaxhps_amd as jh
    from jaxhps_amd import workloads as wl

    print("config D small", flush=True)
    dom, kw, g = wl.config_D(L=1, p=8, nsrc=4)
    pde = jh.PDEProblem(dom, **kw)
    jh.build_solver(pde)
    u = _np(jh.solve(pde, g))
    half = 2.0 / 4
    ops = hps3d.leaf_ops_3D(8, 6, half)
    names = {"D_xx_coefficients": "D_xx", "D_yy_coefficients": "D_yy", "D_zz_coefficients": "D_zz",
             "D_x_coefficients": "D_x", "D_y_coefficients": "D_y", "D_z_coefficients": "D_z"}
    A = hps2d.assemble_diff_operator({names[k]: v for k, v in kw.items() if k in names}, ops, order=ORDER3)
    Y, T, v, h = hps2d.leaf_dtn(A, kw["source"], ops["P"], ops["Q"])
    S, gl, _, _ = hps3d.merge_stage_3d(T, h, 1)
    u0, _ = hps3d.down_pass_3d(g, S, gl, Y, v)
    assert _rel(u, u0) < 1e-10
    # and it approximates the manufactured solution
    assert np.abs(u - wl.exact_D(dom)).max() < 1e-2


@pytest.mark.parametrize("p,L", [(8, 1), (10, 2)])
def test_config_d_mixed_precision_matches_oracle(gpu, p, L):
    """Config D's workload (bumps coefficient, 4 right-hand sides) built in
    FP32 (leaves and every merge level, the root included) and refined in
    FP64: within 1e-8 of the oracle's stage-by-stage FP64 solution."""
    import jaxhps_amd as jh
    from jaxhps_amd import workloads as wl

    print(f"config D mixed p={p} L={L}", flush=True)
    dom, kw, g = wl.config_D(L=L, p=p, nsrc=4)
    pde = jh.PDEProblem(dom, **kw)
    jh.build_solver(pde, precision="mixed")
    assert str(pde._factors["leaf"].lu.dtype) == "torch.float32"
    assert all(str(S.dtype) == "torch.float32" for S in pde.S_lst)
    u = jh.solve(pde, g, refine_tol=1e-12)
    assert str(u.dtype) == "torch.float64"
    assert pde.refine_converged, pde.refine_history
    u = _np(u)
    ops = hps3d.leaf_ops_3D(p, p - 2, 2.0 / 2 ** (L + 1))
    names = {"D_xx_coefficients": "D_xx", "D_yy_coefficients": "D_yy", "D_zz_coefficients": "D_zz",
             "D_x_coefficients": "D_x", "D_y_coefficients": "D_y", "D_z_coefficients": "D_z"}
    A = hps2d.assemble_diff_operator({names[k]: v for k, v in kw.items() if k in names}, ops, order=ORDER3)
    Y, T, v, h = hps2d.leaf_dtn(A, kw["source"], ops["P"], ops["Q"])
    S, gl, _, _ = hps3d.merge_stage_3d(T, h, L)
    u0, _ = hps3d.down_pass_3d(g, S, gl, Y, v)
    assert np.linalg.norm(u - u0) / np.linalg.norm(u0) < 1e-8
