"""Up pass for a new source with a prebuilt solver, named like the
reference's ``up_pass`` package (``up_pass/_uniform_2D_ItI.py:8-219``,
``up_pass/_uniform_2D_DtN.py:8-173``).

The reference stores explicit inverses (Phi, D^{-1}, B D^{-1}); this package
keeps the LU factors of every leaf and merge matrix (and B, D12, D21) from
the build and applies them: v = (leaf operator)^{-1} f, h = Q v, then per
level g_tilde = -D^{-1} h_int and h_ext' = h_ext + B g_tilde.
"""

from __future__ import annotations

import torch

from . import _lib, profiling
from ._stages import cast, leaf_apply_inverse, to_device, up_level


def _factors(pde_problem):
    f = getattr(pde_problem, "_factors", None)
    if f is None:
        raise ValueError("the solver was built without factors; call build_solver first")
    return f


def _up(source, pde_problem, return_h_last=False, inj_lst=None, rhs_full=False, hinj=None):
    """v and g_tilde_lst for `source` through the stored factors.

    rhs_full=True takes `source` as the full leaf right-hand side (ItI: both
    the boundary rows of B = [G; A_int] and the interior rows) instead of a
    PDE source, and returns 3D outputs.  hinj (nl, nq, nsrc) is added to the
    leaves' outgoing data h (the DtN interface residual of refinement).  Each
    merge level runs in the precision its factors were built in."""
    fac = _factors(pde_problem)
    leaf = fac["leaf"]
    dev, dt = leaf.lu.device, leaf.lu.dtype
    src = to_device(source, dev, dt)
    multi = src.ndim == 3 or rhs_full
    if src.ndim == 2:
        src = src.unsqueeze(-1)
    if rhs_full:
        rhs = src.contiguous()
    else:
        nb = pde_problem.P.shape[0]
        rhs = src.clone()
        if pde_problem.use_ItI:
            rhs[:, :nb, :] = 0  # ItI: the boundary rows carry the incoming data, zero here
    with profiling.stage("up_leaf"):
        v = leaf_apply_inverse(leaf, rhs)
        Q = pde_problem.device_ops(dev, dt)["Q"]
        if hinj is not None:
            h = _lib.gemm(Q, v, C=cast(hinj, dt), beta=1.0)
        else:
            h = _lib.gemm(Q, v)
    g_lst = []
    levels = fac["levels"]
    for i, lvl in enumerate(levels):
        last = i == len(levels) - 1
        inj = inj_lst[i] if inj_lst is not None else None
        with profiling.stage(f"up_level_{i + 1}"):
            h = cast(h, lvl.F_lu.dtype)
            gt, h = up_level(h.reshape(-1, lvl.tables.nec, h.shape[-1]), lvl, inj=inj,
                             want_h=(not last) or return_h_last)
        g_lst.append(gt if multi else gt[..., 0])
    v_out = v if multi else v[..., 0]
    if return_h_last:
        return v_out, g_lst, (h[0] if multi else h[0, :, 0])
    return v_out, g_lst


def up_pass_uniform_2D_ItI(source, pde_problem, device=None, host_device=None, return_h_last: bool = False):
    """v and g_tilde_lst for a new source (n_leaves, p^2[, nsrc]) with an ItI
    solver built by build_solver (``up_pass/_uniform_2D_ItI.py:8-136``)."""
    return _up(source, pde_problem, return_h_last)


def up_pass_uniform_2D_DtN(source, pde_problem, device=None, host_device=None, return_h_last: bool = False):
    """DtN version (``up_pass/_uniform_2D_DtN.py:8-107``)."""
    return _up(source, pde_problem, return_h_last)
