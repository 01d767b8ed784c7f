"""Down pass, named and shaped like the reference's ``down_pass`` package.

One ``hps_down_level`` call per level (g_int = S g + g_tilde fused with the
scatter to the children), then one ``hps_leaf_apply`` (u = Y g + v).
"""

from __future__ import annotations

from typing import List

import torch

from ._stages import down_level, leaf_apply, to_device
from ._tables import quad_merge_tables


def _down(boundary_data, S_lst, g_tilde_lst, Y_arr, v_arr, table_fn, name):
    if isinstance(boundary_data, (list, tuple)):
        boundary_data = torch.cat([to_device(b, S_lst[0].device, S_lst[0].dtype) for b in boundary_data])
    S0 = S_lst[-1]
    dev, dt = S0.device, S0.dtype
    bd = to_device(boundary_data, dev, dt)
    multi = len(g_tilde_lst) > 0 and g_tilde_lst[0].ndim == 3
    if multi and bd.ndim == 1:
        raise ValueError("For multi-source downward pass, need to specify boundary data for each source.")
    g = bd.reshape(1, bd.shape[0], -1) if bd.ndim == 2 else bd.reshape(1, -1, 1)
    if not multi and g.shape[-1] != 1:
        multi = True
    for lvl in range(len(S_lst) - 1, -1, -1):
        S = S_lst[lvl].contiguous()
        gt = g_tilde_lst[lvl] if g_tilde_lst else None  # empty after a no-source build
        if gt is not None:
            gt = gt if gt.ndim == 3 else gt.unsqueeze(-1)
            if gt.shape[-1] != g.shape[-1]:
                gt = gt.expand(-1, -1, g.shape[-1])
            gt = gt.contiguous()
        tb = table_fn(S.shape[-1])
        g = down_level(S, gt, g.contiguous(), tb)
    if Y_arr is None:
        return g if multi else g[..., 0]
    v = v_arr
    if v is not None:
        v = v if v.ndim == 3 else v.unsqueeze(-1)
        if v.shape[-1] != g.shape[-1]:
            v = v.expand(-1, -1, g.shape[-1])
        v = v.contiguous()
    u = leaf_apply(Y_arr.contiguous(), g, v)
    return u if multi else u[..., 0]


def down_pass_uniform_2D_DtN(boundary_data, S_lst, g_tilde_lst, Y_arr, v_arr, device=None, host_device=None):
    """Dirichlet data on the root boundary -> leaf solutions
    (``down_pass/_uniform_2D_DtN.py:7-122``)."""
    return _down(
        boundary_data, S_lst, g_tilde_lst, Y_arr, v_arr, lambda ne: quad_merge_tables(ne // 8, "dtn"), "dtn"
    )


def down_pass_uniform_2D_ItI(boundary_data, S_lst, g_tilde_lst, Y_arr, v_arr, device=None, host_device=None):
    """Incoming impedance data on the root boundary -> leaf solutions
    (``down_pass/_uniform_2D_ItI.py:8-118``)."""
    return _down(
        boundary_data, S_lst, g_tilde_lst, Y_arr, v_arr, lambda ne: quad_merge_tables(ne // 8, "iti"), "iti"
    )
