"""Down pass, named and shaped like the reference's ``down_pass`` package.

One ``hps_down_level`` call per level (g_int = S g + g_tilde fused with the
scatter to the children), then one ``hps_leaf_apply`` (u = Y g + v).
"""

from __future__ import annotations

from typing import List

import torch

from ._stages import cast, down_level, leaf_apply, to_device
from ._tables import oct_merge_tables, quad_merge_tables


def down_levels(g_root, S_lst, g_tilde_lst, Y_arr, v_arr, table_fn=None):
    """Root data (ne_root, nsrc) -> (u (nl, n, nsrc) or None, leaf boundary data
    (nl, nec, nsrc)).  Each level runs in the precision of its S; the leaf
    apply in the precision of Y.  g_tilde_lst may be empty (no-source build).
    table_fn(ne) gives a level's merge tables (2D quad tables by default)."""
    g = g_root.reshape(1, g_root.shape[0], -1)
    for lvl in range(len(S_lst) - 1, -1, -1):
        S = S_lst[lvl].contiguous()
        g = cast(g, S.dtype)
        gt = g_tilde_lst[lvl] if g_tilde_lst else None
        if gt is not None:
            gt = gt if gt.ndim == 3 else gt.unsqueeze(-1)
            if gt.shape[-1] != g.shape[-1]:
                gt = gt.expand(-1, -1, g.shape[-1])
            gt = cast(gt.contiguous(), S.dtype)
        tb = table_fn(S.shape[-1]) if table_fn is not None else _tables_for(S)
        g = down_level(S, gt, g.contiguous(), tb)
    if Y_arr is None:
        return None, g
    g = cast(g, Y_arr.dtype)
    v = v_arr
    if v is not None:
        v = v if v.ndim == 3 else v.unsqueeze(-1)
        if v.shape[-1] != g.shape[-1]:
            v = v.expand(-1, -1, g.shape[-1])
        v = cast(v.contiguous(), Y_arr.dtype)
    return leaf_apply(Y_arr.contiguous(), g, v), g


def _tables_for(S):
    ni, ne = S.shape[-2], S.shape[-1]
    kind = "iti" if ni == ne else "dtn"  # ItI S is 8m x 8m, DtN S is 4m x 8m
    return quad_merge_tables(ne // 8, kind)


def _down(boundary_data, S_lst, g_tilde_lst, Y_arr, v_arr, return_leaf_g=False, table_fn=None):
    if isinstance(boundary_data, (list, tuple)):
        boundary_data = torch.cat([to_device(b, S_lst[0].device, S_lst[-1].dtype) for b in boundary_data])
    S0 = S_lst[-1]
    dev, dt = S0.device, S0.dtype
    bd = to_device(boundary_data, dev, dt)
    multi = len(g_tilde_lst) > 0 and g_tilde_lst[0].ndim == 3
    if multi and bd.ndim == 1:
        raise ValueError("For multi-source downward pass, need to specify boundary data for each source.")
    g = bd.reshape(bd.shape[0], -1)
    if not multi and g.shape[-1] != 1:
        multi = True
    u, g_leaf = down_levels(g, S_lst, g_tilde_lst, Y_arr, v_arr, table_fn=table_fn)
    out = g_leaf if u is None else u
    out = out if multi else out[..., 0]
    if return_leaf_g:
        return out, g_leaf
    return out


def down_pass_uniform_2D_DtN(boundary_data, S_lst, g_tilde_lst, Y_arr, v_arr, device=None, host_device=None):
    """Dirichlet data on the root boundary -> leaf solutions
    (``down_pass/_uniform_2D_DtN.py:7-122``)."""
    return _down(boundary_data, S_lst, g_tilde_lst, Y_arr, v_arr)


def down_pass_uniform_2D_ItI(boundary_data, S_lst, g_tilde_lst, Y_arr, v_arr, device=None, host_device=None):
    """Incoming impedance data on the root boundary -> leaf solutions
    (``down_pass/_uniform_2D_ItI.py:8-118``)."""
    return _down(boundary_data, S_lst, g_tilde_lst, Y_arr, v_arr)


def down_pass_uniform_3D_DtN(boundary_data, S_lst, g_tilde_lst, Y_arr, v_arr, device=None, host_device=None):
    """Dirichlet data on the root boundary (faces 1..6) -> leaf solutions
    (``down_pass/_uniform_3D_DtN.py:8-113``).  The root S and g_tilde come
    without a batch axis, as merge_stage_uniform_3D_DtN stores them."""
    S_lst = list(S_lst)
    g_tilde_lst = list(g_tilde_lst)
    if S_lst and S_lst[-1].ndim == 2:
        S_lst[-1] = S_lst[-1].unsqueeze(0)
        if g_tilde_lst:
            g_tilde_lst[-1] = g_tilde_lst[-1].unsqueeze(0)
    return _down(boundary_data, S_lst, g_tilde_lst, Y_arr, v_arr, table_fn=lambda ne: oct_merge_tables(ne // 24))


def down_pass_adaptive_2D_DtN(pde_problem, boundary_data):
    """Adaptive 2D down pass (``down_pass/_adaptive_2D_DtN.py:16-98``)."""
    from ._adaptive import down_pass_adaptive_2D_DtN as fn

    return fn(pde_problem, boundary_data)


def down_pass_adaptive_3D_DtN(pde_problem, boundary_data):
    """Adaptive 3D down pass (``down_pass/_adaptive_3D_DtN.py:15-142``)."""
    from ._adaptive3d import down_pass_adaptive_3D_DtN as fn

    return fn(pde_problem, boundary_data)
