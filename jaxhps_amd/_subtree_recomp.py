"""Subtree recomputation (``_subtree_recomp.py:20-541`` of the reference):
solve a uniform 2D problem by splitting the tree into subtrees of height
``subtree_height``, keeping only the subtree roots' maps on the way up and
recomputing each subtree's factors on the way down.  Nothing below the
subtree roots is stored, so trees deeper than the stored-factor build allows
fit in HBM.
"""

from __future__ import annotations

import copy

import torch

from ._stages import to_device
from .down_pass import down_pass_uniform_2D_DtN, down_pass_uniform_2D_ItI
from .local_solve import local_solve_stage_uniform_2D_DtN, local_solve_stage_uniform_2D_ItI
from .merge import merge_stage_uniform_2D_DtN, merge_stage_uniform_2D_ItI

_COEFS = ("D_xx", "D_xy", "D_yy", "D_xz", "D_yz", "D_zz", "D_x", "D_y", "D_z", "I")


def _slice(a, s, e):
    return None if a is None else a[s:e]


def pde_problem_chunk(pde_problem, start: int, end: int, L: int):
    """A PDEProblem for leaves [start, end) (a subtree of height L) sharing the
    parent's operators (``_pdeproblem.py:336-420`` does the same)."""
    from ._pdeproblem import PDEProblem

    new = PDEProblem.__new__(PDEProblem)
    new.__dict__.update({k: v for k, v in pde_problem.__dict__.items() if not k.endswith("_coefficients")})
    for nm in _COEFS:
        setattr(new, f"{nm}_coefficients", _slice(getattr(pde_problem, f"{nm}_coefficients", None), start, end))
    new.source = _slice(pde_problem.source, start, end)
    dom = copy.copy(pde_problem.domain)
    dom.L = L
    dom.n_leaves = end - start
    dom.interior_points = _slice(pde_problem.domain.interior_points, start, end)
    new.domain = dom
    # operators only: the refinement caches hold whole-tree data
    new._dev_cache = {k: v for k, v in pde_problem._dev_cache.items() if not str(k[0]).startswith("refine")}
    new.reset()
    return new


def _fns(pde_problem):
    if pde_problem.use_ItI:
        return local_solve_stage_uniform_2D_ItI, merge_stage_uniform_2D_ItI, down_pass_uniform_2D_ItI
    return local_solve_stage_uniform_2D_DtN, merge_stage_uniform_2D_DtN, down_pass_uniform_2D_DtN


def _check(pde_problem):
    if not pde_problem.domain.bool_2D:
        raise ValueError("Subtree recomputation is only supported for 2D problems.")
    if not pde_problem.domain.bool_uniform:
        raise ValueError("Subtree recomputation is only supported for uniform quadtrees.")


def _local_solve_and_build(pde_problem, boundary_data, subtree_height, return_top_T=False):
    """Upward pass (boundary_data None): the subtree roots' maps and the top
    merges; downward pass: every subtree rebuilt and solved with its slice of
    boundary_data (the subtree roots' boundary data)."""
    local_solve, merge, down = _fns(pde_problem)
    dom = pde_problem.domain
    chunk = 4**subtree_height
    n_chunks = dom.n_leaves // chunk
    T_lst, h_lst, u_lst = [], [], []
    for i in range(n_chunks):
        sub = pde_problem_chunk(pde_problem, i * chunk, (i + 1) * chunk, subtree_height)
        Y, T, v, h = local_solve(sub)
        if boundary_data is None:
            T_top, h_top = merge(T, h, subtree_height, subtree_recomp=True)
            T_lst.append(T_top)
            h_lst.append(h_top)
        else:
            S_lst, g_lst = merge(T, h, subtree_height)
            u_lst.append(down(boundary_data[i], S_lst, g_lst, Y, v))
        del Y, T, v, h
    if boundary_data is None:
        return merge(torch.cat(T_lst), torch.cat(h_lst), dom.L - subtree_height, return_T=return_top_T)
    return torch.cat(u_lst)


def _all_together(pde_problem, boundary_data):
    local_solve, merge, down = _fns(pde_problem)
    Y, T, v, h = local_solve(pde_problem)
    S_lst, g_lst = merge(T, h, pde_problem.domain.L)
    return down(boundary_data, S_lst, g_lst, Y, v)


def solve_subtree(pde_problem, boundary_data, subtree_height: int = 7, compute_device=None, host_device=None):
    """Solve with subtree recomputation (``_subtree_recomp.py:20-166``); no
    solver state is kept.  Returns u (n_leaves, p^2[, nsrc])."""
    _check(pde_problem)
    if isinstance(boundary_data, (list, tuple)):
        boundary_data = torch.cat([to_device(b, torch.device("cpu"), torch.as_tensor(b).dtype) for b in boundary_data])
    if 4**subtree_height >= pde_problem.domain.n_leaves:
        return _all_together(pde_problem, boundary_data)
    S_lst, g_lst = _local_solve_and_build(pde_problem, None, subtree_height)
    _, _, down = _fns(pde_problem)
    bdry = down(boundary_data, S_lst, g_lst, None, None)
    return _local_solve_and_build(pde_problem, bdry, subtree_height)


def upward_pass_subtree(pde_problem, subtree_height: int = 7, compute_device=None, host_device=None):
    """Upward pass of subtree recomputation (``_subtree_recomp.py:395-455``):
    stores the top levels' S_lst and g_tilde_lst in pde_problem and returns the
    top-level map."""
    _check(pde_problem)
    S_lst, g_lst, T = _local_solve_and_build(pde_problem, None, subtree_height, return_top_T=True)
    pde_problem.S_lst = S_lst
    pde_problem.g_tilde_lst = g_lst
    return T


def downward_pass_subtree(pde_problem, boundary_data, subtree_height: int = 7, compute_device=None,
                          host_device=None):
    """Downward pass of subtree recomputation (``_subtree_recomp.py:458-541``)."""
    _check(pde_problem)
    if isinstance(boundary_data, (list, tuple)):
        boundary_data = torch.cat([torch.as_tensor(b) for b in boundary_data])
    _, _, down = _fns(pde_problem)
    bdry = down(boundary_data, pde_problem.S_lst, pde_problem.g_tilde_lst, None, None)
    return _local_solve_and_build(pde_problem, bdry, subtree_height)
