"""The learned leaf-ItI surrogate: BF16-MFMA GEMM layouts, MLP gradients
against torch autograd (FP64, CPU), training on the method's own leaf solves,
and the probe check with plain-leaf fallback."""

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _rb(t):
    """Round to bfloat16 and back (what the kernel does to GEMM operands)."""
    return t.to(torch.bfloat16).to(torch.float64)


@pytest.mark.parametrize("ta,tb", [(0, 0), (0, 1), (1, 0), (1, 1)])
@pytest.mark.parametrize("M,N,K", [(70, 90, 100), (16, 33, 32), (130, 64, 257)])
def test_bf16_gemm_layouts(gpu, ta, tb, M, N, K):
    from jaxhps_amd import _lib

    print(f"bf16 gemm ta={ta} tb={tb} {M}x{N}x{K}", flush=True)
    g = torch.Generator().manual_seed(M + N + K + 2 * ta + tb)
    A = torch.randn((K, M) if ta else (M, K), generator=g)
    B = torch.randn((N, K) if tb else (K, N), generator=g)
    C0 = torch.randn((M, N), generator=g)
    bias = torch.randn(N, generator=g)
    lib = _lib.load_library()
    h = _lib.handle()
    Ad, Bd, Cd, bd = A.to(gpu), B.to(gpu), C0.to(gpu), bias.to(gpu)
    for act, beta in [(0, 0.0), (1, 0.5)]:
        Cx = Cd.clone()
        rc = lib.hps_mlp_gemm_bf16(h, ta, tb, M, N, K, _lib.ptr(Ad), A.shape[1], _lib.ptr(Bd), B.shape[1], beta,
                                   _lib.ptr(Cx), N, _lib.ptr(bd), act)
        _lib.check(rc, h, "gemm")
        opA = _rb(A).T if ta else _rb(A)
        opB = _rb(B).T if tb else _rb(B)
        ref = opA @ opB + beta * C0.double() + bias.double()
        if act:
            ref = torch.tanh(ref)
        err = (Cx.cpu().double() - ref).abs().max().item()
        assert err < 1e-5 * max(1.0, ref.abs().max().item()) * K**0.5, err


def test_mlp_gradients_match_torch_autograd(gpu):
    from jaxhps_amd.surrogate import LeafItISurrogate

    print("mlp gradients", flush=True)
    model = LeafItISurrogate(n_in=32, width=64, nq=8, rank=2, seed=3, device=gpu)
    g = torch.Generator().manual_seed(5)
    X = torch.randn((16, 32), generator=g, dtype=torch.float32)
    target = torch.complex(torch.randn((16, 8, 8), generator=g), torch.randn((16, 8, 8), generator=g))
    Xd = X.to(gpu)
    H1, H2, Y = model.forward(Xd)
    loss, dY = model.loss_and_grad(Y, target.to(gpu))
    model.backward(Xd, H1, H2, dY)
    # FP64 torch reference of the same network and loss
    P = {k: v.detach().cpu().double().requires_grad_(True) for k, v in model.params.items()}
    h1 = torch.tanh(X.double() @ P["W1"].T + P["b1"])
    h2 = torch.tanh(h1 @ P["W2"].T + P["b2"])
    y = h2 @ P["W3"].T + P["b3"]
    ur = 8 * 2
    U = torch.complex(y[:, :ur], y[:, ur:2 * ur]).reshape(16, 8, 2)
    V = torch.complex(y[:, 2 * ur:3 * ur], y[:, 3 * ur:]).reshape(16, 8, 2)
    E = U @ V.conj().transpose(1, 2) - target.to(torch.complex128)
    ref_loss = (E.abs() ** 2).sum() / 16
    ref_loss.backward()
    assert abs(float(loss) - float(ref_loss.detach())) < 2e-2 * float(ref_loss.detach())
    for k in ["W1", "b1", "W2", "b2", "W3", "b3"]:
        ours = model.grads[k].cpu().double()
        ref = P[k].grad
        rel = (ours - ref).norm() / ref.norm()
        assert rel < 3e-2, (k, float(rel))


def test_surrogate_training_gain_and_fallback(gpu):
    """Trained on FP64 leaf solves of random sine-basis fields (k = 40, p = 16,
    leaves of a level-5 tree), the network's maps are much closer to the exact
    leaf ItI maps of held-out fields than the uncorrected q = 0 map; the probe
    check keeps surrogate leaves only below its tolerance."""
    from jaxhps_amd.surrogate import LeafItISurrogate, leaf_training_data, surrogate_leaf_maps

    print("surrogate training", flush=True)
    q, T, T0 = leaf_training_data(n_fields=8, L=5, p=16, k=40.0, seed=1, device=gpu)
    n_train = 6 * 1024
    model = LeafItISurrogate(n_in=256, width=512, nq=56, rank=8, seed=0, device=gpu)
    hist = model.fit(q[:n_train], T[:n_train], T0, steps=1200, batch=256, lr=1e-3, log_every=200)
    assert hist[-1][1] < 0.5 * hist[0][1], hist
    qt, Tt = q[n_train:], T[n_train:]
    Ts = model.predict_T(qt)
    nrm = Tt.abs().pow(2).sum((1, 2)).sqrt()
    err_sur = ((Ts - Tt).abs().pow(2).sum((1, 2)).sqrt() / nrm).mean().item()
    err_t0 = ((T0[None] - Tt).abs().pow(2).sum((1, 2)).sqrt() / nrm).mean().item()
    print(f"held-out relative error: surrogate {err_sur:.3e}, q=0 map {err_t0:.3e}", flush=True)
    assert err_sur < 0.5 * err_t0
    chk = surrogate_leaf_maps(model, qt, Tt, tol=float(chk_tol := np.median(
        ((Ts - Tt).abs().pow(2).sum((1, 2)).sqrt() / nrm).cpu().numpy())))
    assert 0 < int(chk.passed.sum()) < qt.shape[0]
    assert float(chk.probe_error[chk.passed].max()) <= chk_tol
    fb = ~chk.passed
    assert torch.equal(chk.T[fb], Tt[fb])


def _config_e_problem(gpu, L, seed_theta=0):
    import jaxhps_amd as jh
    from jaxhps_amd import workloads as wl

    k = 40.0
    dom, kw, g, basis, thetas, theta_star = wl.config_E(L=L, p=16, k=k)
    I = k**2 * (1.0 - np.tensordot(thetas[seed_theta], basis, axes=1))
    pde = jh.PDEProblem(dom, I_coefficients=I, **kw)
    return k, dom, kw, g, I, pde


def test_shared_leaf_solver_matches_plain_leaf_solves(gpu):
    """The fixed-point leaf solver with the one shared factorisation of the
    q = 0 leaf matrix equals the plain per-leaf LU solves (config E's leaves:
    k = 40, p = 16, L = 6)."""
    from jaxhps_amd._lib import HPS_KIND_ITI
    from jaxhps_amd._stages import leaf_apply_inverse, leaf_solve
    from jaxhps_amd.surrogate import SharedLeafSolver

    print("shared leaf solver", flush=True)
    k, dom, kw, g, I, pde = _config_e_problem(gpu, 6)
    solver = SharedLeafSolver(pde, I_ref=k**2, device=gpu)
    gen = torch.Generator(device="cpu").manual_seed(2)
    rhs = torch.complex(torch.randn((dom.n_leaves, 256, 2), generator=gen, dtype=torch.float64),
                        torch.randn((dom.n_leaves, 256, 2), generator=gen, dtype=torch.float64)).to(gpu)
    x, conv, it = solver.solve(rhs, torch.as_tensor(I, device=gpu))
    assert bool(conv.all()) and it <= 30, it
    res = leaf_solve(pde, HPS_KIND_ITI, torch.complex128, gpu, nsrc_zero=True)
    x_ref = leaf_apply_inverse(res, rhs)
    assert float((x - x_ref).abs().max() / x_ref.abs().max()) < 1e-10


def test_surrogate_solve_with_probe_check_and_fallback(gpu):
    """The surrogate forward solve (MLP leaf maps, each checked on a random
    boundary probe, complex64 merges, plain interior solves) against the plain
    FP64 build + solve on a held-out field at config E's size (L = 6); with a
    zero tolerance every leaf falls back and the result is the plain one."""
    import jaxhps_amd as jh
    from jaxhps_amd.surrogate import LeafItISurrogate, SharedLeafSolver, leaf_training_data, surrogate_solve

    print("surrogate solve in the loop", flush=True)
    q, T, T0 = leaf_training_data(n_fields=4, L=6, p=16, k=40.0, seed=1, device=gpu)
    model = LeafItISurrogate(n_in=256, width=512, nq=56, rank=8, seed=0, device=gpu)
    model.fit(q, T, T0, steps=1000, batch=256, lr=1e-3)
    k, dom, kw, g, I, pde = _config_e_problem(gpu, 6, seed_theta=3)
    jh.build_solver(pde)
    u_plain = jh.solve(pde, g)
    solver = SharedLeafSolver(pde, I_ref=k**2, device=gpu)
    It = torch.as_tensor(I, device=gpu)
    res = surrogate_solve(model, solver, pde, g, It, k, tol=2e-2)
    frac = float(res.passed.double().mean())
    err = float((res.u - u_plain).norm() / u_plain.norm())
    print(f"surrogate leaves {frac:.3f}, end-to-end error vs plain {err:.2e}, fixed-point steps {res.iterations}",
          flush=True)
    assert frac > 0.5
    assert float(res.probe_error[res.passed].max()) <= 2e-2
    assert err < 0.1
    allplain = surrogate_solve(model, solver, pde, g, It, k, tol=0.0, merge_dtype=torch.complex128)
    assert not bool(allplain.passed.any())
    assert float((allplain.u - u_plain).norm() / u_plain.norm()) < 1e-9
