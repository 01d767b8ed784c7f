"""GPU parity on the benchmark workloads' own generators (jaxhps_amd.workloads)
at tree depths the CPU oracle finishes in seconds: the FP64 path against the
stage-by-stage oracle and against the exact discrete solution (1e-10), the
mixed-precision refined path against both (1e-8), and the config E gradient
at its real p, k and number of sine terms against central differences."""

import numpy as np
import pytest
import torch

import config_oracle as co

pytestmark = pytest.mark.gpu


def _np(t):
    return t.detach().cpu().numpy()


def _run(dom, kw, g, precision, **bkw):
    import jaxhps_amd as jh

    pde = jh.PDEProblem(dom, **kw)
    jh.build_solver(pde, precision=precision, **bkw)
    u = jh.solve(pde, g, refine_tol=1e-12) if precision == "mixed" else jh.solve(pde, g)
    return pde, _np(u)


@pytest.fixture(scope="module")
def helmholtz_k100_L4():
    from jaxhps_amd import workloads as wl

    dom, kw, g = wl.config_C(L=4)
    return dom, kw, g, co.stage_solution(dom, kw, g)[..., 0], co.exact_solution(dom, kw, g)[..., 0]


def test_helmholtz_k100_fp64_matches_oracle(gpu, helmholtz_k100_L4):
    """Config C's workload (k = 100, 20 bumps, p = 16) at L = 4."""
    dom, kw, g, u_stage, u_exact = helmholtz_k100_L4
    print("config C fp64 L=4", flush=True)
    _, u = _run(dom, kw, g, "fp64")
    assert co.rel2(u, u_stage) < 1e-10
    assert co.rel2(u, u_exact) < 1e-10


@pytest.mark.parametrize("top", [0, 2])
def test_helmholtz_k100_mixed_matches_oracle(gpu, helmholtz_k100_L4, top):
    dom, kw, g, u_stage, u_exact = helmholtz_k100_L4
    print(f"config C mixed L=4 fp64_top_levels={top}", flush=True)
    pde, u = _run(dom, kw, g, "mixed", fp64_top_levels=top)
    assert pde.refine_converged, pde.refine_history
    assert co.rel2(u, u_stage) < 1e-8
    assert co.rel2(u, u_exact) < 1e-8


def test_variable_coefficient_16rhs_matches_oracle(gpu):
    """Config B's workload (D_xx = D_yy = -a, D_x, D_y, I = c; 16 right-hand
    sides; p = 16) at L = 3 with a random source: FP64 at 1e-10 and mixed at
    1e-8 against the oracle and the exact discrete solution."""
    from jaxhps_amd import workloads as wl

    print("config B L=3 16 rhs", flush=True)
    dom, kw, g = wl.config_B(L=3)
    kw = dict(kw)
    kw["source"] = np.random.default_rng(5).standard_normal(kw["source"].shape)
    u_stage = co.stage_solution(dom, kw, g)
    u_exact = co.exact_solution(dom, kw, g)
    _, u64 = _run(dom, kw, g, "fp64")
    assert u64.shape == u_stage.shape
    assert co.rel2(u64, u_stage) < 1e-10
    assert co.rel2(u64, u_exact) < 1e-10
    pde, umx = _run(dom, kw, g, "mixed")
    assert pde.refine_converged, pde.refine_history
    assert co.rel2(umx, u_stage) < 1e-8
    assert co.rel2(umx, u_exact) < 1e-8


def _theta_setup(gpu, L):
    import jaxhps_amd as jh
    from jaxhps_amd import workloads as wl

    k = 40.0
    dom, kw, g, basis, thetas, theta_star = wl.config_E(L=L, p=16, k=k)
    basis_t = torch.as_tensor(basis, device=gpu)
    return jh, k, dom, kw, g, basis, basis_t, thetas, theta_star


def test_inverse_scattering_forward_matches_oracle(gpu):
    """Config E's forward problem (k = 40, p = 16, 25 sine terms) at L = 3:
    u(q_theta) against the oracle."""
    jh, k, dom, kw, g, basis, basis_t, thetas, _ = _theta_setup(gpu, 3)
    print("config E forward L=3", flush=True)
    I = k**2 * (1.0 - np.einsum("j,jln->ln", thetas[0], basis))
    kw2 = dict(kw, I_coefficients=I)
    _, u = _run(dom, kw2, g, "fp64")
    u_stage = co.stage_solution(dom, kw2, g)[..., 0]
    assert co.rel2(u, u_stage) < 1e-10


def test_inverse_scattering_gradient_matches_central_fd(gpu):
    """dJ/dtheta of J = ||u(q_theta) - u_obs||^2 at config E's p = 16, k = 40
    and 25 sine terms (L = 3): the adjoint against central differences of
    the FP64 forward solve, 1e-6 relative to the largest component."""
    from jaxhps_amd.adjoint import differentiable_solve

    jh, k, dom, kw, g, basis, basis_t, thetas, theta_star = _theta_setup(gpu, 3)
    print("config E gradient vs FD L=3 25 terms", flush=True)
    flat = basis_t.reshape(basis_t.shape[0], -1)
    shape = basis_t.shape[1:]
    g_dev = torch.as_tensor(g, device=gpu)
    ones = np.ones(dom.interior_points.shape[:2])
    pde = jh.PDEProblem(dom, I_coefficients=ones, **kw)
    with torch.no_grad():
        u_obs = differentiable_solve(pde, g_dev, I=k**2 * (1.0 - (torch.as_tensor(theta_star, device=gpu) @ flat)
                                                          .reshape(shape))).clone()

    def J(theta):
        u = differentiable_solve(pde, g_dev, I=k**2 * (1.0 - (theta @ flat).reshape(shape)))
        return ((u - u_obs).abs() ** 2).sum()

    theta = torch.tensor(thetas[0], device=gpu, requires_grad=True)
    J(theta).backward()
    grad = _np(theta.grad)
    assert grad.shape == (25,)
    eps = 1e-6
    fd = np.zeros(25)
    with torch.no_grad():
        for j in range(25):
            e = torch.zeros(25, dtype=torch.float64, device=gpu)
            e[j] = eps
            fd[j] = (float(J(theta.detach() + e)) - float(J(theta.detach() - e))) / (2 * eps)
    assert np.abs(grad - fd).max() <= 1e-6 * np.abs(fd).max(), (grad, fd)


def test_poisson_config_a_matches_oracle_and_manufactured_solution(gpu):
    """Config A (2D Poisson, p = 8, L = 3, DtN, u = sin(pi x) sin(pi y)):
    FP64 GPU against the oracle (1e-10) and the manufactured solution (the
    discretisation error at p = 8)."""
    from jaxhps_amd import workloads as wl

    print("config A", flush=True)
    dom, kw, g = wl.config_A()
    _, u = _run(dom, kw, g, "fp64")
    u_stage = co.stage_solution(dom, kw, g)[..., 0]
    assert co.rel2(u, u_stage) < 1e-10
    assert co.rel2(u, co.exact_solution(dom, kw, g)[..., 0]) < 1e-10
    assert np.abs(u - wl.exact_A(dom)).max() < 1e-5
