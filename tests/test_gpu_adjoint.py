"""The hand-written adjoint of build+solve: against torch autograd through the
oracle's dense global system (tiny trees), against central finite
differences of the inverse-scattering objective (config E's problem at small
size) and of a real DtN objective."""

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


@pytest.mark.parametrize("L", [1, 2])
def test_iti_gradient_matches_dense_autograd(gpu, L):
    """dJ/dI for J = sum |u - u_obs|^2 against torch autograd through the dense
    ItI system assembled by the oracle from coordinates."""
    import jaxhps_amd as jh
    from jaxhps_amd.adjoint import differentiable_solve

    print(f"adjoint vs dense autograd L={L}", flush=True)
    p, q = 6, 4
    dom = jh.Domain(p=p, q=q, root=jh.DiscretizationNode2D(*cs.BOUNDS), L=L)
    pts = dom.interior_points
    ones = np.ones(pts.shape[:2])
    cI = 2.0 + np.cos(pts[..., 0]) * np.sin(pts[..., 1])
    f = cs.cc_source(pts)
    rng = np.random.default_rng(L)
    t_in = rng.standard_normal(dom.boundary_points.shape[0]) + 1j * rng.standard_normal(dom.boundary_points.shape[0])
    u_obs = rng.standard_normal(pts.shape[:2]) + 1j * rng.standard_normal(pts.shape[:2])
    # ours
    pde = jh.PDEProblem(dom, source=f, D_xx_coefficients=ones, D_yy_coefficients=ones, I_coefficients=cI,
                        use_ItI=True, eta=cs.ETA)
    I_t = torch.tensor(cI, device=gpu, requires_grad=True)
    u = differentiable_solve(pde, t_in, I=I_t)
    J = ((u - torch.as_tensor(u_obs, device=gpu)).abs() ** 2).sum()
    J.backward()
    g_ours = _np(I_t.grad)
    # dense reference
    half = (dom.root.xmax - dom.root.xmin) / 2 ** (L + 1)
    ops = osp.leaf_ops_2D(p, q, half, eta=cs.ETA)
    A0 = hps2d.assemble_diff_operator({"D_xx": ones, "D_yy": ones}, ops)
    M0, rhs, per = gs.iti_system(0 * A0, f, ops["P_iti"], ops["QH"], ops["G"], cs.BOUNDS, L, q, t_in,
                                 dom.boundary_points)
    n = p * p
    nb = ops["P_iti"].shape[0]
    c = torch.tensor(cI, dtype=torch.float64, requires_grad=True)
    A = torch.as_tensor(A0).to(torch.complex128) + torch.diag_embed(c.to(torch.complex128))
    M = torch.as_tensor(M0).clone()
    for l in range(dom.n_leaves):
        M[l * per + nb : l * per + n, l * per : l * per + n] = A[l, nb:, :]
    x = torch.linalg.solve(M, torch.as_tensor(rhs))
    ud = torch.stack([x[l * per : l * per + n] for l in range(dom.n_leaves)])
    assert np.abs(_np(ud) - _np(u)).max() < 1e-10 * np.abs(_np(ud)).max()
    Jd = ((ud - torch.as_tensor(u_obs)).abs() ** 2).sum()
    Jd.backward()
    g_ref = c.grad.numpy()
    assert np.abs(g_ours - g_ref).max() < 1e-9 * np.abs(g_ref).max()
    # coefficients on the leaf boundary rows never enter the system
    assert np.abs(g_ours[:, :nb]).max() == 0.0


def _theta_objective(dom, kw, g, basis_t, u_obs, k, theta):
    import jaxhps_amd as jh
    from jaxhps_amd.adjoint import differentiable_solve

    q = torch.einsum("j,jln->ln", theta, basis_t)
    I = k**2 * (1.0 - q)
    pde = jh.PDEProblem(dom, I_coefficients=np.ones(dom.interior_points.shape[:2]), **kw)
    u = differentiable_solve(pde, g, I=I)
    return ((u - u_obs).abs() ** 2).sum()


def test_sine_theta_gradient_matches_central_fd(gpu):
    """Config E's objective J(theta) = ||u(q_theta) - u_obs||^2 (k = 10, L = 3,
    p = 10): autograd through the adjoint vs central differences."""
    from jaxhps_amd import workloads as wl

    print("adjoint theta FD", flush=True)
    k = 10.0
    dom, kw, g, basis, thetas, theta_star = wl.config_E(L=3, p=10, k=k, n_theta=1, n_terms=6, seed=3)
    basis_t = torch.as_tensor(basis, device=gpu)
    with torch.no_grad():
        import jaxhps_amd as jh

        pde = jh.PDEProblem(dom, I_coefficients=_np(k**2 * (1 - torch.einsum("j,jln->ln", torch.as_tensor(theta_star, device=gpu), basis_t))), **kw)
        jh.build_solver(pde)
        u_obs = jh.solve(pde, g).clone()
    theta = torch.tensor(thetas[0], device=gpu, requires_grad=True)
    J = _theta_objective(dom, kw, g, basis_t, u_obs, k, theta)
    J.backward()
    grad = _np(theta.grad)
    eps = 1e-5
    fd = np.zeros_like(grad)
    with torch.no_grad():
        for j in range(len(grad)):
            e = torch.zeros_like(theta)
            e[j] = eps
            fd[j] = (float(_theta_objective(dom, kw, g, basis_t, u_obs, k, theta.detach() + e))
                     - float(_theta_objective(dom, kw, g, basis_t, u_obs, k, theta.detach() - e))) / (2 * eps)
    assert np.abs(grad - fd).max() <= 1e-6 * np.abs(fd).max(), (grad, fd)


def test_dtn_real_gradient_matches_central_fd(gpu):
    """Real DtN problem (config B's coefficients, small tree): directional
    derivative of J along random perturbations of I and D_xx."""
    import jaxhps_amd as jh
    from jaxhps_amd import workloads as wl
    from jaxhps_amd.adjoint import differentiable_solve

    print("adjoint dtn FD", flush=True)
    dom, kw, g = wl.config_B(L=2, p=10, nsrc=2, seed=4)
    rng = np.random.default_rng(9)
    kw = dict(kw)
    kw["source"] = rng.standard_normal(kw["source"].shape)
    u_obs = torch.as_tensor(rng.standard_normal(kw["source"].shape), device=gpu)
    c0 = torch.as_tensor(kw.pop("I_coefficients"), device=gpu)
    a0 = torch.as_tensor(kw.pop("D_xx_coefficients"), device=gpu)
    dc = torch.as_tensor(rng.standard_normal(c0.shape), device=gpu)
    da = torch.as_tensor(rng.standard_normal(a0.shape), device=gpu) * 0.1

    def J(c, a):
        pde = jh.PDEProblem(dom, I_coefficients=_np(c0), D_xx_coefficients=_np(a0), **kw)
        u = differentiable_solve(pde, g, I=c, D_xx=a)
        return ((u - u_obs) ** 2).sum()

    c = c0.clone().requires_grad_(True)
    a = a0.clone().requires_grad_(True)
    Jv = J(c, a)
    Jv.backward()
    dirderiv = float((c.grad * dc).sum() + (a.grad * da).sum())
    eps = 1e-6
    with torch.no_grad():
        fd = (float(J(c0 + eps * dc, a0 + eps * da)) - float(J(c0 - eps * dc, a0 - eps * da))) / (2 * eps)
    assert abs(dirderiv - fd) <= 1e-6 * abs(fd), (dirderiv, fd)


def test_checkpointed_adjoint_equals_stored_factors(gpu):
    """checkpoint=True drops the merge levels after the forward solve and
    recomputes them in backward: the gradient is the one of the stored-factor
    adjoint (config E's problem at p = 16, k = 40, L = 3, 25 sine terms), and
    no merge level is held between forward and backward."""
    from jaxhps_amd import workloads as wl
    import jaxhps_amd as jh
    from jaxhps_amd.adjoint import differentiable_solve

    print("checkpointed adjoint", flush=True)
    k = 40.0
    dom, kw, g, basis, thetas, theta_star = wl.config_E(L=3, p=16, k=k)
    flat = torch.as_tensor(basis, device=gpu).reshape(basis.shape[0], -1)
    shape = basis.shape[1:]
    g_dev = torch.as_tensor(g, device=gpu)
    u_obs = torch.zeros(shape, dtype=torch.complex128, device=gpu)
    grads = []
    for ckpt in (False, True):
        pde = jh.PDEProblem(dom, I_coefficients=np.ones(shape), **kw)
        theta = torch.tensor(thetas[1], device=gpu, requires_grad=True)
        u = differentiable_solve(pde, g_dev, checkpoint=ckpt, I=k**2 * (1.0 - (theta @ flat).reshape(shape)))
        if ckpt:
            assert pde._factors["levels"] is None and pde.S_lst == []
        ((u - u_obs).abs() ** 2).sum().backward()
        grads.append(_np(theta.grad))
    assert np.abs(grads[1] - grads[0]).max() <= 1e-12 * np.abs(grads[0]).max()
