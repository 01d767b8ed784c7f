"""Hand-written adjoint of build_solver + solve with respect to the
coefficient fields, behind a torch.autograd.Function.

The reference gets gradients from JAX reverse mode through build_solver and
solve (``examples/check_autodiff_vJp.py:154-243``,
``examples/inverse_wave_scattering.py:209-292``), which stores every
intermediate.  Here the uniform build+solve is one linear solve u = K^{-1} b
with the global collocation system K of the build (leaf rows plus interface
coupling), so for a real loss J(u) with torch's incoming gradient G = dJ/du
(torch's complex convention):

    lambda = K^{-H} G,      dJ/dc_k[l, i] = -Re(conj(lambda[l, i]) (D_k u)[l, i])

for the rows i the coefficients enter (interior rows of each leaf; zero on
the leaf boundary rows, which the reference's leaf solves never read).
lambda is the exact adjoint of the solve's own stages, applied in reverse
through the factors the build kept (nothing is refactorised):

  1. gbar_leaf = Y^H G                         (adjoint of u = Y g + v)
  2. leaves -> root: the adjoint of each down-pass level gives the interface
     adjoints gtbar (and S^H for the parent's data)
  3. root -> leaves: the adjoint of each up-pass (merge) level,
     hbar_int = -D^{-H}(gtbar + B^H hbar_out), scattered to the children
  4. vbar = G + Q^H hbar_leaf, lambda = (leaf operator)^{-H} vbar.

By default every factor stays resident in HBM between the forward solve and
the backward pass (288 GB per GPU holds every level's for the configurations
here).  With checkpoint=True the forward pass keeps only the leaf stage's
factors and maps and drops every merge level; the backward pass recomputes
the merge levels from the stored leaf maps (the same deterministic kernels,
so the same factors) and frees them again after the adjoint.
"""

from __future__ import annotations

import ctypes
from typing import Dict

import torch

from . import _lib
from ._lib import HPS_KIND_DTN, HPS_KIND_ITI, check, dtype_code, load_library, ptr, workspace
from ._stages import to_device


def _gemm_conjt(A, lda, sA, M, K, B, ldb, sB, N, batch, C=None, beta=0.0, out=None):
    lib = load_library()
    hd = _lib.handle()
    if out is None:
        out = torch.empty((batch, M, N), dtype=A.dtype, device=A.device)
    rc = lib.hps_gemm_batched_conjt(hd, dtype_code(A.dtype), M, N, K, 1.0, ptr(A), lda, sA, ptr(B), ldb, sB,
                                    float(beta), ptr(C) if C is not None else None, N, M * N if C is not None else 0,
                                    ptr(out), N, M * N, batch)
    check(rc, hd, "hps_gemm_batched_conjt")
    return out


def getrs_conjt(n, lu, lu_offset, lda, sA, perm, dinv, R):
    """X = A^{-H} R from factors stored in `lu` (element offset, row stride lda, batch stride sA)."""
    lib = load_library()
    hd = _lib.handle()
    b, _, nrhs = R.shape
    X = torch.empty_like(R)
    ws = workspace(lib.hps_getrs_conjt_workspace(dtype_code(lu.dtype), n, nrhs, b), R.device)
    es = lu.element_size()
    rc = lib.hps_getrs_batched_conjt(hd, dtype_code(lu.dtype), n, nrhs, ctypes.c_void_p(lu.data_ptr() + lu_offset * es),
                                     lda, sA, ptr(perm), ptr(dinv), ptr(R), nrhs, n * nrhs, ptr(X), nrhs, n * nrhs, b,
                                     ptr(ws), ws.numel())
    check(rc, hd, "hps_getrs_batched_conjt")
    return X


def _adjoint_down_level(gbar_children, S, tables):
    lib = load_library()
    hd = _lib.handle()
    nm = S.shape[0]
    w = gbar_children.shape[-1]
    _, _, ext, _, cols_final, _ = tables.device(S.device)
    gext = torch.empty((nm, tables.ne, w), dtype=S.dtype, device=S.device)
    gt = torch.empty((nm, tables.ni, w), dtype=S.dtype, device=S.device)
    rc = lib.hps_adjoint_down_level(hd, dtype_code(S.dtype), nm, tables.nc, tables.nec, tables.ni, tables.ne, w,
                                    ptr(S.contiguous()), ptr(gbar_children.contiguous()), ptr(ext), ptr(cols_final),
                                    ptr(gext), ptr(gt))
    check(rc, hd, "hps_adjoint_down_level")
    return gext, gt


def _adjoint_up_level(gtbar, hbar_out, lvl):
    lib = load_library()
    hd = _lib.handle()
    tb = lvl.tables
    dev, dt = gtbar.device, gtbar.dtype
    kind = HPS_KIND_ITI if tb.kind == "iti" else HPS_KIND_DTN
    nm = lvl.F_lu.shape[0]
    w = gtbar.shape[-1]
    fin = tb.device(dev)[5]
    adj = tb.adj_device(dev)
    out = torch.empty((nm * tb.nc, tb.nec, w), dtype=dt, device=dev)
    ws = workspace(lib.hps_adjoint_up_workspace(dtype_code(dt), kind, nm, tb.ni, w), dev)
    rc = lib.hps_adjoint_up_level(
        hd, dtype_code(dt), kind, nm, tb.nc, tb.nec, tb.ni, tb.ne, w, ptr(gtbar.contiguous()),
        ptr(hbar_out.contiguous()) if hbar_out is not None else None, ptr(lvl.B), ptr(lvl.F_lu), ptr(lvl.perm),
        ptr(lvl.dinv), ptr(lvl.D12), ptr(lvl.D21), ptr(fin), ptr(adj), ptr(out), ptr(ws), ws.numel(),
    )
    check(rc, hd, "hps_adjoint_up_level")
    return out


def adjoint_lambda(pde_problem, G):
    """lambda = K^{-H} G for the system of an FP64 build_solver; G (n_leaves, n[, nsrc])."""
    fac = getattr(pde_problem, "_factors", None)
    if fac is None:
        raise ValueError("adjoint needs a solver built by build_solver")
    if getattr(pde_problem, "_precision", "fp64") != "fp64":
        raise NotImplementedError("the adjoint runs on the FP64 build")
    leaf, levels = fac["leaf"], fac["levels"]
    dev, dt = leaf.lu.device, leaf.lu.dtype
    squeeze = G.ndim == 2
    Gd = to_device(G, dev, dt)
    if Gd.ndim == 2:
        Gd = Gd.unsqueeze(-1)
    Gd = Gd.contiguous()
    nl, n, w = Gd.shape
    Y = leaf.Y
    nq = Y.shape[-1]
    # 1. gbar_leaf = Y^H G
    gbar = _gemm_conjt(Y, nq, n * nq, nq, n, Gd, w, n * w, w, nl)
    # 2. adjoint down pass, leaves -> root
    gtbars = []
    for lvl in levels:
        gbar, gt = _adjoint_down_level(gbar.reshape(-1, lvl.tables.nec, w), lvl.S, lvl.tables)
        gtbars.append(gt)
    # 3. adjoint up pass, root -> leaves
    hbar = None
    for lvl, gt in zip(reversed(levels), reversed(gtbars)):
        hbar_out = None if hbar is None else hbar.reshape(-1, lvl.tables.ne, w)
        hbar = _adjoint_up_level(gt, hbar_out, lvl)
    # 4. vbar = G + Q^H hbar_leaf ; lambda = (leaf operator)^{-H} vbar
    Q = pde_problem.device_ops(dev, dt)["Q"]
    vbar = _gemm_conjt(Q, n, 0, n, nq, hbar.contiguous(), w, nq * w, w, nl, C=Gd, beta=1.0)
    _, nrows, _ = leaf.lu.shape
    if leaf.kind == HPS_KIND_ITI:
        # the factors are of B' = diag(s, 1) B, so B^{-H} = diag(s, 1) B'^{-H}
        lam = getrs_conjt(n, leaf.lu, 0, n, n * n, leaf.perm, leaf.dinv, vbar)
        if leaf.row_scale is not None:
            lam[:, : leaf.nb, :] *= leaf.row_scale.to(lam.real.dtype)
    else:
        nb = n - nrows
        lam = torch.zeros_like(vbar)
        lam[:, nb:, :] = getrs_conjt(nrows, leaf.lu, nb, n, nrows * n, leaf.perm, leaf.dinv,
                                     vbar[:, nb:, :].contiguous())
    return lam[..., 0] if squeeze else lam


def coefficient_gradients(pde_problem, u, lam) -> Dict[str, torch.Tensor]:
    """dJ/dc for every coefficient field present: -Re(conj(lam) (D_k u)) on the
    rows the coefficients enter.  Returns {name: (n_leaves, n) float64}."""
    from ._refine import refine_ops

    lib = load_library()
    hd = _lib.handle()
    u3 = u if u.ndim == 3 else u.unsqueeze(-1)
    l3 = lam if lam.ndim == 3 else lam.unsqueeze(-1)
    u3 = u3.contiguous()
    l3 = l3.to(u3.dtype).contiguous()
    ops = refine_ops(pde_problem, u3.device, u3.dtype)
    nl, n, w = u3.shape
    R = ops.products(u3)
    nb = pde_problem.P.shape[0]
    ncoef = len(ops.coef_ops)
    grad = torch.empty((ncoef, nl, n), dtype=torch.float64, device=u3.device)
    rc = lib.hps_coef_gradient(hd, dtype_code(u3.dtype), nl, n, nb, w, ops.nops, ncoef,
                               ops.coef_ops.ctypes.data_as(ctypes.c_void_p), ptr(R), ops.nrow_R,
                               ptr(u3), ptr(l3), ptr(grad))
    check(rc, hd, "hps_coef_gradient")
    names = [nm for nm, _ in pde_problem.coefficient_list()]
    return {nm: grad[k] for k, nm in enumerate(names)}


class _BuildSolve(torch.autograd.Function):
    """u = solve(build_solver(pde with the given coefficient fields), g)."""

    @staticmethod
    def forward(ctx, pde, boundary_data, names, checkpoint, *coefs):
        from ._build_solver import build_solver
        from ._solve import solve

        pde.reset()  # drops the factors and the cached coefficient copies
        for nm, c in zip(names, coefs):
            setattr(pde, f"{nm}_coefficients", c.detach())
        build_solver(pde)
        u = solve(pde, boundary_data)
        if checkpoint:
            # keep the leaf factors and maps; the merge levels are recomputed in backward
            pde._factors["levels"] = None
            pde.S_lst = []
            pde.g_tilde_lst = []
        ctx.pde = pde
        ctx.names = names
        ctx.checkpoint = checkpoint
        ctx.save_for_backward(u)
        return u

    @staticmethod
    def backward(ctx, grad_u):
        (u,) = ctx.saved_tensors
        pde = ctx.pde
        if ctx.checkpoint:
            from ._build_solver import _merge_levels

            leaf = pde._factors["leaf"]
            levels, _ = _merge_levels(pde, leaf.T, leaf.h, False)
            pde._factors["levels"] = levels
        lam = adjoint_lambda(pde, grad_u)
        grads = coefficient_gradients(pde, u, lam)
        if ctx.checkpoint:
            pde._factors["levels"] = None
        return (None, None, None, None) + tuple(grads[nm] for nm in ctx.names)


def differentiable_solve(pde_problem, boundary_data, checkpoint: bool = False, **coefficients):
    """Build the solver with the given coefficient fields (e.g. I=tensor)
    and solve; the result is differentiable with respect to those fields
    through the hand-written adjoint.  Coefficient tensors must be real
    (n_leaves, p^d) tensors; other fields come from pde_problem.
    checkpoint=True keeps only the leaf factors between forward and backward
    and recomputes the merge levels in backward (less memory, one more merge
    stage per gradient)."""
    names = tuple(coefficients)
    for nm in names:
        if getattr(pde_problem, f"{nm}_coefficients", None) is None and coefficients[nm] is None:
            raise ValueError(f"no {nm} coefficients")
        if coefficients[nm].is_complex():
            raise ValueError("differentiable coefficient fields must be real")
    coefs = [coefficients[nm] for nm in names]
    return _BuildSolve.apply(pde_problem, boundary_data, names, bool(checkpoint), *coefs)
