"""Device-side stage drivers: one C-ABI call per leaf stage / tree level.

These are the only functions that talk to libhps_mi355x; the public stage
functions (named like the reference's) and build_solver/solve sit on top.
All arrays are torch tensors resident in HBM.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import List, Optional

import numpy as np
import torch

from . import _lib
from ._lib import HPS_KIND_DTN, HPS_KIND_ITI, check, dtype_code, load_library, ptr, workspace
from ._tables import MergeTables


def default_device():
    _lib._require_gpu()
    return torch.device("cuda", torch.cuda.current_device())


def to_device(x, device, dtype):
    if x is None:
        return None
    if isinstance(x, torch.Tensor):
        t = x
    else:
        t = torch.as_tensor(np.asarray(x))
    if not t.is_complex() and dtype.is_complex:
        t = t.to(torch.float64)
    return t.to(device=device, dtype=dtype).contiguous()


def is_complex_input(*arrs) -> bool:
    for a in arrs:
        if a is None:
            continue
        if isinstance(a, torch.Tensor):
            if a.is_complex():
                return True
        elif np.iscomplexobj(a):
            return True
    return False


def real_dtype(dt):
    return {torch.complex128: torch.float64, torch.complex64: torch.float32}.get(dt, dt)


def complex_dtype(dt):
    return {torch.float64: torch.complex128, torch.float32: torch.complex64}.get(dt, dt)


@dataclass
class LeafResult:
    Y: torch.Tensor
    T: torch.Tensor
    v: Optional[torch.Tensor]
    h: Optional[torch.Tensor]
    lu: torch.Tensor
    perm: torch.Tensor
    info: torch.Tensor
    kind: int


def leaf_solve(pde, kind: int, dtype, device, source=None, nsrc_zero=False) -> LeafResult:
    """Run hps_leaf_solve over all leaves of `pde`.

    source: (n_leaves, n, nsrc) tensor in `dtype`, or None for a zero source
    (nsrc_zero=True gives the no-source leaf stage, nsrc = 0)."""
    lib = load_library()
    h = _lib.handle()
    dom = pde.domain
    nl = dom.n_leaves
    n = dom.p**2 if dom.bool_2D else dom.p**3
    ops = pde.device_ops(device, dtype)
    P, Q = ops["P"], ops["Q"]
    nb, nq = P.shape
    names = pde.operator_names()
    clist = pde.coefficient_list()
    if not clist:
        raise ValueError("PDEProblem has no coefficients")
    coef_ops = np.array([-1 if nm == "I" else names.index(nm) for nm, _ in clist], dtype=np.int32)
    coeffs = torch.stack([to_device(a, device, dtype) for _, a in clist]).contiguous()
    if nsrc_zero:
        nsrc = 0
        src = None
    else:
        if source is None:
            nsrc = 1
            src = None
        else:
            nsrc = source.shape[-1]
            src = source
    nrows = n - nb if kind == HPS_KIND_DTN else n
    Y = torch.empty((nl, n, nq), dtype=dtype, device=device)
    T = torch.empty((nl, nq, nq), dtype=dtype, device=device)
    v = torch.empty((nl, n, nsrc), dtype=dtype, device=device) if nsrc else None
    hv = torch.empty((nl, nq, nsrc), dtype=dtype, device=device) if nsrc else None
    lu = torch.empty((nl, nrows, n), dtype=dtype, device=device)
    perm = torch.empty((nl, nrows), dtype=torch.int32, device=device)
    info = torch.empty((nl,), dtype=torch.int32, device=device)
    ws = workspace(lib.hps_leaf_workspace(dtype_code(dtype), kind, nl, n, nb, nq, nsrc), device)
    cop = coef_ops.ctypes.data_as(ctypes.c_void_p)
    rc = lib.hps_leaf_solve(
        h, dtype_code(dtype), kind, nl, n, nb, nq, nsrc, len(clist), cop,
        ptr(coeffs), ptr(ops["diffops"]), ptr(P), ptr(Q), ptr(ops.get("G")), ptr(src),
        ptr(Y), ptr(T), ptr(v), ptr(hv), ptr(lu), ptr(perm), ptr(info), ptr(ws), ws.numel(),
    )
    check(rc, h, "hps_leaf_solve")
    return LeafResult(Y, T, v, hv, lu, perm, info, kind)


@dataclass
class MergeLevel:
    S: torch.Tensor
    T: Optional[torch.Tensor]
    h: Optional[torch.Tensor]
    gt: Optional[torch.Tensor]
    D_lu: torch.Tensor
    perm: torch.Tensor
    info: torch.Tensor
    B: Optional[torch.Tensor]
    tables: MergeTables


def merge_level(T_in, h_in, tables: MergeTables, add_identity: bool, keep_B=False, want_T=True) -> MergeLevel:
    """One level of 4:1 / 8:1 merges.  T_in (nm*nc, nec, nec), h_in (nm*nc, nec, nsrc) or None."""
    lib = load_library()
    hd = _lib.handle()
    dev = T_in.device
    dt = T_in.dtype
    nc, nec, ni, ne = tables.nc, tables.nec, tables.ni, tables.ne
    nm = T_in.shape[0] // nc
    assert T_in.shape == (nm * nc, nec, nec), (T_in.shape, nm, nc, nec)
    nsrc = 0 if h_in is None else h_in.shape[-1]
    rows, cols, ext, _ = tables.device(dev)
    S = torch.empty((nm, ni, ne), dtype=dt, device=dev)
    T = torch.empty((nm, ne, ne), dtype=dt, device=dev) if want_T else None
    ho = torch.empty((nm, ne, nsrc), dtype=dt, device=dev) if nsrc else None
    gt = torch.empty((nm, ni, nsrc), dtype=dt, device=dev) if nsrc else None
    D_lu = torch.empty((nm, ni, ni), dtype=dt, device=dev)
    perm = torch.empty((nm, ni), dtype=torch.int32, device=dev)
    info = torch.empty((nm,), dtype=torch.int32, device=dev)
    B = torch.empty((nm, ne, ni), dtype=dt, device=dev) if keep_B else None
    ws = workspace(lib.hps_merge_workspace(dtype_code(dt), nm, ni, ne, nsrc), dev)
    rc = lib.hps_merge_level(
        hd, dtype_code(dt), nm, nc, nec, ni, ne, nsrc, int(add_identity),
        ptr(T_in), ptr(h_in), ptr(rows), ptr(cols), ptr(ext),
        ptr(S), ptr(T), ptr(ho), ptr(gt), ptr(D_lu), ptr(perm), ptr(info), ptr(B), ptr(ws), ws.numel(),
    )
    check(rc, hd, "hps_merge_level")
    return MergeLevel(S, T, ho, gt, D_lu, perm, info, B, tables)


def down_level(S, gt, g_in, tables: MergeTables):
    """g_in (nm, ne, nsrc) -> children boundary data (nm*nc, nec, nsrc)."""
    lib = load_library()
    hd = _lib.handle()
    dev = S.device
    dt = S.dtype
    nc, nec, ni, ne = tables.nc, tables.nec, tables.ni, tables.ne
    nm = S.shape[0]
    nsrc = g_in.shape[-1]
    _, _, _, src = tables.device(dev)
    g_out = torch.empty((nm * nc, nec, nsrc), dtype=dt, device=dev)
    ws = workspace(lib.hps_down_workspace(dtype_code(dt), nm, ni, nsrc), dev)
    rc = lib.hps_down_level(
        hd, dtype_code(dt), nm, nc, nec, ni, ne, nsrc, ptr(S), ptr(gt), ptr(g_in), ptr(src), ptr(g_out),
        ptr(ws), ws.numel(),
    )
    check(rc, hd, "hps_down_level")
    return g_out


def leaf_apply(Y, g, v):
    """u = Y g + v per leaf."""
    lib = load_library()
    hd = _lib.handle()
    nl, n, nq = Y.shape
    nsrc = g.shape[-1]
    u = torch.empty((nl, n, nsrc), dtype=Y.dtype, device=Y.device)
    rc = lib.hps_leaf_apply(hd, dtype_code(Y.dtype), nl, n, nq, nsrc, ptr(Y), ptr(g), ptr(v), ptr(u))
    check(rc, hd, "hps_leaf_apply")
    return u
