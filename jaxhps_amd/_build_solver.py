"""``build_solver`` (``_build_solver.py:38-174``), uniform trees, device-resident."""

from __future__ import annotations

import torch

from ._lib import raise_on_info
from .local_solve import (
    _leaf,
    local_solve_stage_uniform_2D_DtN,
)
from ._lib import HPS_KIND_DTN, HPS_KIND_ITI
from .merge import merge_stage_uniform_2D_DtN, merge_stage_uniform_2D_ItI


def build_solver(pde_problem, return_top_T: bool = False, compute_device=None, host_device=None):
    """Leaf solves and merges for the whole tree; stores Y, v, S_lst and
    g_tilde_lst in ``pde_problem``.  Returns the top-level T when
    ``return_top_T``.  ``compute_device`` may name the GPU (torch device);
    ``host_device`` is accepted for signature compatibility -- the solver
    state stays in HBM."""
    dom = pde_problem.domain
    if not dom.bool_uniform:
        raise NotImplementedError("adaptive trees are outside this package's scope")
    if pde_problem.source is None:
        raise NotImplementedError("the no-source build (solve(..., source=...)) is not implemented yet")
    if not dom.bool_2D:
        raise NotImplementedError("3D build is not implemented yet")
    from . import profiling

    kind = HPS_KIND_ITI if pde_problem.use_ItI else HPS_KIND_DTN
    with profiling.stage("leaf"):
        res, v, h = _leaf(pde_problem, kind, compute_device)
    raise_on_info(res.info, "leaf stage")
    pde_problem.Y = res.Y
    pde_problem.v = v
    merge_fn = merge_stage_uniform_2D_ItI if pde_problem.use_ItI else merge_stage_uniform_2D_DtN
    out = merge_fn(res.T, h, l=dom.L, return_T=return_top_T)
    pde_problem.S_lst = out[0]
    pde_problem.g_tilde_lst = out[1]
    if return_top_T:
        return out[2]
    return None
