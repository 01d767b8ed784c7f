"""Leaf stage, named and shaped like the reference's ``local_solve`` package.

Each function runs one batched C-ABI call (``hps_leaf_solve``) over all
leaves: fused operator assembly, LU with partial pivoting, triangular solves
and the GEMMs for Y, T, v, h.  Outputs are device tensors.
``device``/``host_device`` are accepted for signature compatibility; results
stay in HBM (on ``device`` when it is a torch device).
"""

from __future__ import annotations

import torch

from ._lib import HPS_KIND_DTN, HPS_KIND_ITI
from ._stages import complex_dtype, default_device, is_complex_input, leaf_solve, to_device


def _resolve(device):
    if isinstance(device, (torch.device, str)) and torch.device(device).type == "cuda":
        return torch.device(device)
    return default_device()


def _compute_dtype(pde, base=torch.float64, force_complex=False):
    arrs = [a for _, a in pde.coefficient_list()] + [pde.source]
    dt = base
    if force_complex or is_complex_input(*arrs):
        dt = complex_dtype(base)
    return dt


def _source_3d(pde, dtype, device):
    src = pde.source
    if src is None:
        return None, False
    multi = src.ndim == 3
    t = to_device(src, device, dtype)
    if not multi:
        t = t.unsqueeze(-1)
    return t.contiguous(), multi


def _leaf(pde, kind, device, dtype=None, squeeze=True):
    device = _resolve(device)
    if dtype is None:
        dtype = _compute_dtype(pde, force_complex=(kind == HPS_KIND_ITI))
    src, multi = _source_3d(pde, dtype, device)
    res = leaf_solve(pde, kind, dtype, device, source=src)
    v, h = res.v, res.h
    if squeeze and not multi:
        v = v[..., 0]
        h = h[..., 0]
    return res, v, h


def local_solve_stage_uniform_2D_DtN(pde_problem, host_device=None, device=None):
    """DtN leaf stage (``local_solve/_uniform_2D_DtN.py:9-102``).

    Returns Y (n_leaves, p^2, 4q), T (n_leaves, 4q, 4q), v (n_leaves, p^2[, nsrc]),
    h (n_leaves, 4q[, nsrc])."""
    res, v, h = _leaf(pde_problem, HPS_KIND_DTN, device)
    return res.Y, res.T, v, h


def local_solve_stage_uniform_3D_DtN(pde_problem, host_device=None, device=None):
    """3D DtN leaf stage (``local_solve/_uniform_3D_DtN.py:13-106``); same kernel,
    3D operators (P maps 6q^2 Gauss to p^3-(p-2)^3 Chebyshev boundary points)."""
    res, v, h = _leaf(pde_problem, HPS_KIND_DTN, device)
    return res.Y, res.T, v, h


def local_solve_stage_uniform_2D_ItI(pde_problem, device=None, host_device=None):
    """ItI leaf stage (``local_solve/_uniform_2D_ItI.py:10-117``).

    Returns (Y, T, v, h) in the reference's order; complex128."""
    res, v, h = _leaf(pde_problem, HPS_KIND_ITI, device)
    return res.Y, res.T, v, h


def local_solve_stage_adaptive_2D_DtN(pde_problem, device=None, host_device=None):
    """Adaptive 2D leaf stage (``local_solve/_adaptive_2D_DtN.py:13-117``)."""
    from ._adaptive import local_solve_stage_adaptive_2D_DtN as fn

    return fn(pde_problem, device=device, host_device=host_device)


def local_solve_stage_adaptive_3D_DtN(pde_problem, device=None, host_device=None):
    """Adaptive 3D leaf stage (``local_solve/_adaptive_3D_DtN.py:13-128``)."""
    from ._adaptive3d import local_solve_stage_adaptive_3D_DtN as fn

    return fn(pde_problem, device=device, host_device=host_device)
