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
    dinv: torch.Tensor
    info: torch.Tensor
    kind: int
    # ItI: the factored matrix is diag(s I_nb, I) B -- the incoming-impedance
    # rows scaled by the power of two s (a 0-dim device tensor), see
    # iti_row_scale; 1 for DtN.
    row_scale: Optional[torch.Tensor] = None
    nb: int = 0


def iti_row_scale(pde, coeffs, clist, device):
    """Power-of-two factor s for the impedance rows G of the ItI leaf matrix
    B = [G; A_int], as a 0-dim float64 device tensor.

    G's entries are O(p^2/h) while the collocation rows carry O((p^2/h)^2)
    differentiation entries, so B is badly row-scaled: at p = 16 its
    condition number is ~3e5, and ~6e3 once the G rows are brought to the
    size of the A rows.  Factoring diag(s, 1) B instead of B (with the same
    scaling of the right-hand side's G rows) leaves every solution unchanged
    and cuts the factorisation's rounding error about tenfold.  s estimates
    max|A_int| / max|G| from the coefficient maxima and the operators'
    maxima; a power of two keeps the scaling exact."""
    key = ("iti_row_scale_ops", tuple(nm for nm, _ in clist))
    opmax = pde._dev_cache.get(key)
    if opmax is None:
        opmax = [1.0 if nm == "I" else float(np.abs(np.asarray(getattr(pde, nm))).max()) for nm, _ in clist]
        gmax = float(np.abs(np.asarray(pde.G)).max())
        opmax = pde._dev_cache[key] = (torch.tensor(opmax, dtype=torch.float64, device=device), gmax)
    om, gmax = opmax
    cmax = torch.amax(coeffs.abs().reshape(coeffs.shape[0], -1), dim=1).to(torch.float64)
    est = torch.amax(cmax * om.to(cmax.device)) / gmax
    return torch.exp2(torch.round(torch.log2(torch.clamp(est, min=1e-300))))


def scale_rhs_rows(leaf, rhs):
    """rhs (n_leaves, n, nsrc) with its first nb rows (the impedance rows)
    multiplied by the ItI leaf's row scale, as a new contiguous tensor."""
    out = rhs.contiguous().clone()
    if leaf.kind == HPS_KIND_ITI and leaf.row_scale is not None:
        out[:, : leaf.nb, :] *= leaf.row_scale.to(out.real.dtype if out.is_complex() else out.dtype)
    return out


def scale_batched(X, scale):
    """X[b] *= scale[b] in place (X contiguous, scale a float64 device vector)."""
    lib = load_library()
    h = _lib.handle()
    b = X.shape[0]
    n = X[0].numel()
    rc = lib.hps_scale_batched(h, dtype_code(X.dtype), b, n, n, ptr(scale), ptr(X))
    check(rc, h, "hps_scale_batched")


def leaf_solve(pde, kind: int, dtype, device, source=None, nsrc_zero=False, leaf_half_sides=None) -> LeafResult:
    """Run hps_leaf_solve over all leaves of `pde`.

    source: (n_leaves, n, nsrc) tensor in `dtype`, or None for a zero source
    (nsrc_zero=True gives the no-source leaf stage, nsrc = 0).
    leaf_half_sides (n_leaves,) float64 numpy array: leaves of different
    sizes with operators built for half side 1 -- each coefficient of a
    k-th order term is scaled by half^-k and T, h by 1/half (the operators
    are homogeneous in the half side)."""
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
    if leaf_half_sides is not None:
        hs = np.asarray(leaf_half_sides, dtype=np.float64)
        inv = torch.as_tensor(1.0 / hs).to(device)
        inv2 = torch.as_tensor(1.0 / (hs * hs)).to(device)
        for k, (nm, _) in enumerate(clist):
            order = 0 if nm == "I" else len(nm) - 2
            if order:
                scale_batched(coeffs[k], inv if order == 1 else inv2)
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
    dinv = torch.empty((nl, _lib.dinv_size(nrows)), dtype=dtype, device=device)
    info = torch.empty((nl,), dtype=torch.int32, device=device)
    ws = workspace(lib.hps_leaf_workspace(dtype_code(dtype), kind, nl, n, nb, nq, nsrc), device)
    cop = coef_ops.ctypes.data_as(ctypes.c_void_p)
    G = ops.get("G")
    row_scale = None
    if kind == HPS_KIND_ITI:
        row_scale = iti_row_scale(pde, coeffs, clist, device)
        rs = row_scale.to(real_dtype(dtype))
        G = (G * rs).contiguous()
        P = (P * rs).contiguous()

    def run():
        rc = lib.hps_leaf_solve(
            h, dtype_code(dtype), kind, nl, n, nb, nq, nsrc, len(clist), cop,
            ptr(coeffs), ptr(ops["diffops"]), ptr(P), ptr(Q), ptr(G), ptr(src),
            ptr(Y), ptr(T), ptr(v), ptr(hv), ptr(lu), ptr(perm), ptr(dinv), ptr(info), ptr(ws), ws.numel(),
        )
        check(rc, h, "hps_leaf_solve")

    # natural-order factorisation (the leaf matrices eliminate with multipliers
    # below 2), checked per leaf in the panel kernels: a rejected leaf
    # (info = -1) reruns the whole stage with partial pivoting
    lib.hps_set_leaf_pivoting(h, 0)
    try:
        run()
    finally:
        lib.hps_set_leaf_pivoting(h, 1)
    if bool((info < 0).any()):
        run()
    if leaf_half_sides is not None:
        scale_batched(T, inv)
        if hv is not None:
            scale_batched(hv, inv)
    return LeafResult(Y, T, v, hv, lu, perm, dinv, info, kind, row_scale, nb)


@dataclass
class MergeLevel:
    S: torch.Tensor
    T: Optional[torch.Tensor]
    h: Optional[torch.Tensor]
    gt: Optional[torch.Tensor]
    F_lu: torch.Tensor          # LU of D (DtN) or of W = I - D12 D21 (ItI)
    perm: torch.Tensor
    dinv: torch.Tensor
    info: torch.Tensor
    D12: Optional[torch.Tensor]
    D21: Optional[torch.Tensor]
    B: Optional[torch.Tensor]
    tables: MergeTables
    Finv: Optional[torch.Tensor] = None  # block inverses for few-column up-pass solves (hps_block_inverses)


def merge_level(T_in, h_in, tables: MergeTables, keep_factors=False, want_T=True, pivot=True) -> MergeLevel:
    """One level of 4:1 / 8:1 merges.  T_in (nm*nc, nec, nec), h_in (nm*nc, nec, nsrc) or None.
    keep_factors also keeps B (and D12/D21 for ItI) for the up pass.
    pivot=False factors the interface matrices without row exchanges (see
    merge_factors_stable for the check the caller owes)."""
    lib = load_library()
    hd = _lib.handle()
    dev = T_in.device
    dt = T_in.dtype
    nc, nec, ni, ne = tables.nc, tables.nec, tables.ni, tables.ne
    kind = HPS_KIND_ITI if tables.kind == "iti" else HPS_KIND_DTN
    nm = T_in.shape[0] // nc
    assert T_in.shape == (nm * nc, nec, nec), (T_in.shape, nm, nc, nec)
    nsrc = 0 if h_in is None else h_in.shape[-1]
    nf = ni // 2 if kind == HPS_KIND_ITI else ni
    rows, cols, ext, _, cols_final, final_of_internal = tables.device(dev)
    S = torch.empty((nm, ni, ne), dtype=dt, device=dev)
    T = torch.empty((nm, ne, ne), dtype=dt, device=dev) if want_T else None
    ho = torch.empty((nm, ne, nsrc), dtype=dt, device=dev) if (nsrc and want_T) else None
    gt = torch.empty((nm, ni, nsrc), dtype=dt, device=dev) if nsrc else None
    F_lu = torch.empty((nm, nf, nf), dtype=dt, device=dev)
    perm = torch.empty((nm, nf), dtype=torch.int32, device=dev)
    dinv = torch.empty((nm, _lib.dinv_size(nf)), dtype=dt, device=dev)
    info = torch.empty((nm,), dtype=torch.int32, device=dev)
    keep12 = keep_factors and kind == HPS_KIND_ITI
    D12 = torch.empty((nm, nf, nf), dtype=dt, device=dev) if keep12 else None
    D21 = torch.empty((nm, nf, nf), dtype=dt, device=dev) if keep12 else None
    B = torch.empty((nm, ne, ni), dtype=dt, device=dev) if keep_factors else None
    # the block inverses of low-precision levels for the refinement's up passes
    n_inv = lib.hps_block_inverses_elems(nf) if (dt in (torch.complex64, torch.float32) and nm < 128) else 0
    Finv = torch.empty((nm * n_inv,), dtype=dt, device=dev) if n_inv else None
    ws = workspace(lib.hps_merge_workspace(dtype_code(dt), kind, nm, ni, ne, nsrc), dev)
    upd = np.ascontiguousarray(tables.upd, dtype=np.int32)
    d12 = np.ascontiguousarray(tables.d12_blk, dtype=np.int32)
    d21 = np.ascontiguousarray(tables.d21_blk, dtype=np.int32)
    if not pivot:
        check(lib.hps_set_merge_pivoting(hd, 0), hd, "hps_set_merge_pivoting")
    try:
        rc = _merge_level_call(lib, hd, dt, kind, nm, nc, nec, ni, ne, nsrc, T_in, h_in, rows, cols, ext, cols_final,
                               final_of_internal, upd, d12, d21, S, T, ho, gt, F_lu, perm, dinv, info, D12, D21, B,
                               Finv, ws)
    finally:
        if not pivot:
            lib.hps_set_merge_pivoting(hd, 1)
    check(rc, hd, "hps_merge_level")
    return MergeLevel(S, T, ho, gt, F_lu, perm, dinv, info, D12, D21, B, tables, Finv)


def form_block_inverses(lvl: MergeLevel) -> None:
    """Store in lvl.Finv the inverses of the 256 x 256 diagonal blocks of the
    level's low-precision factors when its up-pass solves would otherwise take
    the blocked few-column path (fewer than 128 matrices of more than 256
    unknowns): the refinement's up passes then solve through them.
    merge_level keeps them already; this forms them for factors from
    elsewhere."""
    nm, n = lvl.F_lu.shape[0], lvl.F_lu.shape[-1]
    if lvl.F_lu.dtype not in (torch.complex64, torch.float32) or nm >= 128:
        return
    lib = load_library()
    elems = lib.hps_block_inverses_elems(n)
    if elems == 0:
        return
    hd = _lib.handle()
    dt = lvl.F_lu.dtype
    inv = torch.empty(nm * elems, dtype=dt, device=lvl.F_lu.device)
    ws = workspace(lib.hps_block_inverses_workspace(dtype_code(dt), n, nm), lvl.F_lu.device)
    rc = lib.hps_block_inverses(hd, dtype_code(dt), n, nm, ptr(lvl.F_lu), ptr(lvl.dinv), ptr(inv), ptr(ws), ws.numel())
    check(rc, hd, "hps_block_inverses")
    lvl.Finv = inv


def merge_factors_stable(F_lu, bound=64.0):
    """Device bool: the natural-order (unpivoted) LU factors of a merge level
    are usable -- every entry finite, no zero pivot, every multiplier at most
    `bound` in magnitude (partial pivoting would guarantee 1; the merge
    matrices of uniform trees give values below 4)."""
    a = F_lu.abs()
    finite = torch.isfinite(a.amax(dim=(-2, -1)))
    mult = torch.tril(a, -1).amax(dim=(-2, -1))
    dmin = a.diagonal(dim1=-2, dim2=-1).amin(dim=-1)
    return (finite & (mult <= bound) & (dmin > 0)).all()


def _merge_level_call(lib, hd, dt, kind, nm, nc, nec, ni, ne, nsrc, T_in, h_in, rows, cols, ext, cols_final,
                      final_of_internal, upd, d12, d21, S, T, ho, gt, F_lu, perm, dinv, info, D12, D21, B, Finv, ws):
    return lib.hps_merge_level(
        hd, dtype_code(dt), kind, nm, nc, nec, ni, ne, nsrc,
        ptr(T_in), ptr(h_in), ptr(rows), ptr(cols), ptr(ext), ptr(cols_final), ptr(final_of_internal),
        upd.shape[0], upd.ctypes.data_as(ctypes.c_void_p),
        d12.shape[0], d12.ctypes.data_as(ctypes.c_void_p), d21.shape[0], d21.ctypes.data_as(ctypes.c_void_p),
        ptr(S), ptr(T), ptr(ho), ptr(gt), ptr(F_lu), ptr(perm), ptr(dinv), ptr(info),
        ptr(D12), ptr(D21), ptr(B), ptr(Finv), ptr(ws), ws.numel(),
    )


def down_level(S, gt, g_in, tables: MergeTables):
    """g_in (nm, ne, nsrc) -> children boundary data (nm*nc, nec, nsrc)."""
    lib = load_library()
    hd = _lib.handle()
    dev = S.device
    dt = S.dtype
    nc, nec, ni, ne = tables.nc, tables.nec, tables.ni, tables.ne
    nm = S.shape[0]
    nsrc = g_in.shape[-1]
    src = tables.device(dev)[3]
    g_out = torch.empty((nm * nc, nec, nsrc), dtype=dt, device=dev)
    ws = workspace(lib.hps_down_workspace(dtype_code(dt), nm, ni, nsrc), dev)
    rc = lib.hps_down_level(
        hd, dtype_code(dt), nm, nc, nec, ni, ne, nsrc, ptr(S), ptr(gt), ptr(g_in), ptr(src), ptr(g_out),
        ptr(ws), ws.numel(),
    )
    check(rc, hd, "hps_down_level")
    return g_out


def down_level_static(S, gt, g_in, tables: MergeTables):
    """down_level through hps_down_level_static: no workspace is allocated
    (the interface data live in the library's static per-GPU scratch).
    Calls that overlap on one GPU race on that scratch."""
    lib = load_library()
    hd = _lib.handle()
    dev = S.device
    dt = S.dtype
    nc, nec, ni, ne = tables.nc, tables.nec, tables.ni, tables.ne
    nm = S.shape[0]
    nsrc = g_in.shape[-1]
    src = tables.device(dev)[3]
    g_out = torch.empty((nm * nc, nec, nsrc), dtype=dt, device=dev)
    rc = lib.hps_down_level_static(hd, dtype_code(dt), nm, nc, nec, ni, ne, nsrc, ptr(S), ptr(gt), ptr(g_in),
                                   ptr(src), ptr(g_out))
    check(rc, hd, "hps_down_level_static")
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


def getrs_into(n, lu, lu_offset, lda, sA, perm, dinv, R, X):
    """X = A^{-1} R from factors stored in `lu` (element offset lu_offset,
    row stride lda, batch stride sA).  R, X: (batch, n, nrhs) contiguous."""
    lib = load_library()
    hd = _lib.handle()
    b = R.shape[0]
    nrhs = R.shape[2]
    es = lu.element_size()
    rc = lib.hps_getrs_batched(
        hd, dtype_code(lu.dtype), n, nrhs, ctypes.c_void_p(lu.data_ptr() + lu_offset * es), lda, sA, ptr(perm),
        ptr(dinv), ptr(R), nrhs, n * nrhs, ptr(X), nrhs, n * nrhs, b,
    )
    check(rc, hd, "hps_getrs_batched")
    return X


def leaf_apply_inverse(leaf: LeafResult, rhs):
    """v = (leaf operator)^{-1} rhs for a full collocation right-hand side
    rhs (n_leaves, n, nsrc): DtN solves the interior rows with A_ii (v is zero
    on the boundary); ItI solves with B = [G; A_int]."""
    nl, nrows, n = leaf.lu.shape
    nsrc = rhs.shape[-1]
    if leaf.kind == HPS_KIND_DTN:
        nb = n - nrows
        R = rhs[:, nb:, :].contiguous()
        X = torch.empty_like(R)
        getrs_into(nrows, leaf.lu, nb, n, nrows * n, leaf.perm, leaf.dinv, R, X)
        v = torch.zeros((nl, n, nsrc), dtype=rhs.dtype, device=rhs.device)
        v[:, nb:, :] = X
        return v
    R = scale_rhs_rows(leaf, rhs)
    X = torch.empty_like(R)
    getrs_into(n, leaf.lu, 0, n, n * n, leaf.perm, leaf.dinv, R, X)
    return X


def up_level(h_in, lvl: MergeLevel, inj=None, want_h=True):
    """New outgoing data through the factors kept by a merge level.
    h_in (nm*nc, nec, nsrc) -> (g_tilde (nm, ni, nsrc) final order, h_out (nm, ne, nsrc) or None)."""
    lib = load_library()
    hd = _lib.handle()
    tb = lvl.tables
    dev, dt = h_in.device, h_in.dtype
    kind = HPS_KIND_ITI if tb.kind == "iti" else HPS_KIND_DTN
    nm = lvl.F_lu.shape[0]
    nsrc = h_in.shape[-1]
    rows, _, ext, _, _, final_of_internal = tb.device(dev)
    gt = torch.empty((nm, tb.ni, nsrc), dtype=dt, device=dev)
    ho = torch.empty((nm, tb.ne, nsrc), dtype=dt, device=dev) if want_h else None
    ws = workspace(lib.hps_up_workspace(dtype_code(dt), kind, nm, tb.ni, nsrc), dev)
    upd = np.ascontiguousarray(tb.upd, dtype=np.int32)
    rc = lib.hps_up_level(
        hd, dtype_code(dt), kind, nm, tb.nc, tb.nec, tb.ni, tb.ne, nsrc, ptr(h_in.contiguous()),
        ptr(inj.contiguous()) if inj is not None else None, ptr(rows), ptr(ext), ptr(final_of_internal),
        upd.shape[0], upd.ctypes.data_as(ctypes.c_void_p), ptr(lvl.F_lu), ptr(lvl.perm), ptr(lvl.dinv),
        ptr(lvl.Finv) if lvl.Finv is not None else None, ptr(lvl.D12), ptr(lvl.D21), ptr(lvl.B), ptr(gt), ptr(ho),
        ptr(ws), ws.numel(),
    )
    check(rc, hd, "hps_up_level")
    return gt, ho


def cast(t, dtype):
    """t converted to `dtype` on the device by hps_convert (t itself if it already has it)."""
    if t is None or t.dtype == dtype:
        return t
    lib = load_library()
    hd = _lib.handle()
    src = t.contiguous()
    out = torch.empty(src.shape, dtype=dtype, device=src.device)
    rc = lib.hps_convert(hd, dtype_code(src.dtype), dtype_code(dtype), src.numel(), ptr(src), ptr(out))
    check(rc, hd, "hps_convert")
    return out
