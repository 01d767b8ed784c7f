"""Merge stage, named and shaped like the reference's ``merge`` package.

Every tree level is one ``hps_merge_level`` call (batched over the level's
merges); the index tables in ``_tables`` carry the reference's block
structure, rolls, flips and row permutation.
"""

from __future__ import annotations

from typing import List

import torch

from ._lib import raise_on_info
from ._stages import cast, merge_factors_stable, merge_level
from ._tables import oct_merge_tables, quad_merge_tables


def _levels(T_arr, h_arr, l, kind, nc, table_fn, keep_B=False, root_T=True, dtypes=None, pivot=False):
    """Run max(l-1, 0) batched levels plus the root merge, as the reference's
    merge stages do (``merge/_uniform_2D_DtN.py:90-203``).  root_T=False skips
    forming the root's T and h (only needed when the caller asks for them).
    dtypes[i], when given, is the precision level i merges in (the children's
    T and h are converted on entry).

    The interface matrices are factored without row exchanges (pivot=False):
    they eliminate stably in their natural order, and pivoted panels are
    latency-bound.  Every level is checked at the end (one synchronisation);
    from the first level that fails the check the merges are redone with
    partial pivoting.  Low-precision levels of fewer than 128 matrices of more
    than 256 unknowns keep the inverses of their factors' diagonal blocks for
    the refinement's few-column up passes (MergeLevel.Finv)."""
    from . import profiling
    nec = T_arr.shape[-1]
    T0 = T_arr.reshape(-1, nec, nec)
    # the reference's multi-source test (merge/_uniform_2D_DtN.py:73)
    multi = h_arr is not None and h_arr.ndim == T_arr.ndim
    if h_arr is None:
        h0 = None
    else:
        h0 = h_arr.reshape(T0.shape[0], nec, -1) if multi else h_arr.reshape(T0.shape[0], nec, 1)
    n_levels = max(l - 1, 0) + 1

    def run(start, T, h, out, piv):
        for lvl in range(start, n_levels):
            if lvl == n_levels - 1:
                T = T[:nc]
                h = h[:nc] if h is not None else None
            if dtypes is not None:
                T = cast(T, dtypes[lvl])
                h = cast(h, dtypes[lvl])
            tb = table_fn(T.shape[-1])
            want_T = root_T or lvl < n_levels - 1
            with profiling.stage(f"merge_level_{lvl + 1}"):
                res = merge_level(T.contiguous(), h.contiguous() if h is not None else None, tb,
                                  keep_factors=keep_B, want_T=want_T, pivot=piv)
            out.append(res)
            T, h = res.T, res.h
        return out

    out = run(0, T0, h0, [], pivot)
    if not pivot:
        # every level was factored without row exchanges: the register-sized
        # panels check themselves (info = -1 on a rejected matrix); the tall
        # panels of the levels with panels taller than the register panels
        # (1024 rows, 512 for complex128) or of at most 16 matrices are
        # checked here on the factors
        tall = lambda r: (r.F_lu.shape[-1] > (512 if r.F_lu.dtype == torch.complex128 else 1024)  # noqa: E731
                          or r.F_lu.shape[0] <= 16)
        ok = torch.stack([(r.info >= 0).all() & (merge_factors_stable(r.F_lu) if tall(r) else True)
                          for r in out]).cpu()
        if not bool(ok.all()):
            first = int((~ok).nonzero()[0, 0])
            T, h = (T0, h0) if first == 0 else (out[first - 1].T, out[first - 1].h)
            out = run(first, T, h, out[:first], True)
    return out, multi


def _pack_outputs(levels, multi, return_T, what):
    for r in levels:
        raise_on_info(r.info, what)
    S_lst = [r.S for r in levels]
    if multi:
        g_lst = [r.gt for r in levels]
    else:
        g_lst = [r.gt[..., 0] for r in levels]
    if return_T:
        return S_lst, g_lst, levels[-1].T[0]
    return S_lst, g_lst


def merge_stage_uniform_2D_DtN(
    T_arr, h_arr, l: int, device=None, host_device=None, subtree_recomp: bool = False, return_T: bool = False
):
    """DtN 4:1 merges (``merge/_uniform_2D_DtN.py:13-203``).

    Returns S_lst (level j: (4^(l-1-j), 4m, 8m); root (1, 4m, 8m)),
    g_tilde_lst and optionally the top T.  subtree_recomp=True returns
    (T_top[None], h_top[None]) like the reference."""
    levels, multi = _levels(T_arr, h_arr, l, "dtn", 4, lambda nec: quad_merge_tables(nec // 4, "dtn"),
                            root_T=return_T or subtree_recomp)
    if subtree_recomp:
        top = levels[-1]
        hh = top.h if multi else top.h[..., 0]
        return top.T, hh
    return _pack_outputs(levels, multi, return_T, "merge_stage_uniform_2D_DtN")


def merge_stage_uniform_2D_ItI(
    T_arr,
    h_arr,
    l: int,
    device=None,
    host_device=None,
    subtree_recomp: bool = False,
    return_T: bool = False,
    return_h: bool = False,
):
    """ItI 4:1 merges (``merge/_uniform_2D_ItI.py:19-179``).  S rows and g_tilde
    come out in the reference's order a5, b5, b6, c6, c7, d7, d8, a8."""
    levels, multi = _levels(T_arr, h_arr, l, "iti", 4, lambda nec: quad_merge_tables(nec // 4, "iti"),
                            root_T=return_T or return_h or subtree_recomp)
    if subtree_recomp:
        top = levels[-1]
        hh = top.h if multi else top.h[..., 0]
        return top.T, hh
    out = _pack_outputs(levels, multi, return_T, "merge_stage_uniform_2D_ItI")
    if return_h:
        top = levels[-1]
        out = tuple(out) + ((top.h[0] if multi else top.h[0, :, 0]),)
    return out


def merge_stage_uniform_3D_DtN(T_arr, h_arr, l: int, device=None, host_device=None, return_T: bool = False):
    """3D DtN 8:1 merges (``merge/_uniform_3D_DtN.py:12-124``).

    Returns S_lst, g_tilde_lst (and the top T with return_T); interface
    unknowns in the reference's order 9..20, exterior faces 1..6 with the
    children's face blocks in the reference's order.  The root level is
    stored without a batch axis, as the reference does (:115-117)."""
    levels, multi = _levels(T_arr, h_arr, l, "dtn", 8, lambda nec: oct_merge_tables(nec // 6), root_T=return_T)
    out = _pack_outputs(levels, multi, return_T, "merge_stage_uniform_3D_DtN")
    S_lst, g_lst = list(out[0]), list(out[1])
    S_lst[-1] = S_lst[-1][0]
    g_lst[-1] = g_lst[-1][0]
    if return_T:
        return S_lst, g_lst, out[2]
    return S_lst, g_lst


def merge_stage_adaptive_2D_DtN(pde_problem) -> None:
    """Adaptive 2D merges (``merge/_adaptive_2D_DtN.py:25-110``); expects
    each leaf's data.T and data.h to be set."""
    from ._adaptive import merge_stage_adaptive_2D_DtN as fn

    fn(pde_problem)


def merge_stage_adaptive_3D_DtN(pde_problem) -> None:
    """Adaptive 3D merges (``merge/_adaptive_3D_DtN.py:30-149``); expects each
    leaf's data.T and data.h to be set."""
    from ._adaptive3d import merge_stage_adaptive_3D_DtN as fn

    fn(pde_problem)
