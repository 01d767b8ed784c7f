"""Mixed-precision solve with FP64 iterative refinement (uniform 2D and 3D trees).

``build_solver(..., precision="mixed")`` factors the leaves and the lower merge
levels in FP32 (complex64 for ItI) and the top ``fp64_top_levels`` merge
levels in FP64.  ``solve`` then computes a first solution through those
factors and refines it: the residual of the global collocation system is
evaluated in FP64 (``hps_residual_{iti,dtn}``), the correction goes through
the same stored factors (up pass + down pass), and the FP64 iterate is
updated until the relative update falls below ``refine_tol``.

The global system is the one the reference's uniform build+solve eliminates
exactly (leaf rows of ``local_solve/_uniform_2D_ItI.py:120-193`` /
``_uniform_2D_DtN.py:228-280``; interface coupling of
``merge/_uniform_2D_ItI.py:183-375`` / ``merge/_schur_complement.py:183-237``),
so the refined solution converges to the FP64 HPS solution.

ItI state: u on every leaf's Chebyshev grid.  Residual rows: [P t_in - G u;
f_int - A_int u] with t_in the root data or minus the neighbour's outgoing
impedance.  The correction is an up pass whose leaf right-hand side has both
row blocks, and a down pass from zero root data.

DtN state (2D and 3D): the leaves' Gauss boundary data g_leaf (shared values
agree on both sides of an interface) and u, whose boundary rows are P g_leaf.  Residuals:
interior rows f - A u, the flux mismatch Q u_a + Q u_b on interior Gauss nodes
(added to one leaf's h in the up pass) and g_root - g_leaf on the root
boundary (the down pass's root data).
"""

from __future__ import annotations

import ctypes
import functools
import warnings

import numpy as np
import torch

from . import _lib, profiling
from ._lib import HPS_KIND_DTN, check, dtype_code, load_library, ptr
from ._neighbors import gauss_neighbors_2D, gauss_neighbors_3D
from ._stages import cast, leaf_apply_inverse, to_device

LOW = {torch.float64: torch.float32, torch.complex128: torch.complex64}
HIGH = {torch.float32: torch.float64, torch.complex64: torch.complex128,
        torch.float64: torch.float64, torch.complex128: torch.complex128}
_NPARTS = 1024


class RefineOps:
    """FP64 operators of the global system for one PDEProblem on one device."""

    def __init__(self, pde, device, dtype):
        dom = pde.domain
        names = pde.operator_names()
        clist = pde.coefficient_list()
        used = sorted({names.index(nm) for nm, _ in clist if nm != "I"})
        self.coef_ops = np.array([-1 if nm == "I" else used.index(names.index(nm)) for nm, _ in clist],
                                 dtype=np.int32)
        self.nops = len(used)
        mats = [np.asarray(getattr(pde, names[i])) for i in used]
        mats += [np.asarray(pde.G), np.asarray(pde.QH)] if pde.use_ItI else [np.asarray(pde.Q)]
        M = np.concatenate(mats, axis=0)
        self.nrow_R = M.shape[0]

        def dev(x):
            return torch.as_tensor(np.ascontiguousarray(x)).to(device=device, dtype=dtype)

        self.M = dev(M)
        self.MT = dev(M.T)
        self.P = dev(pde.P)
        self.coeffs = None  # set by refine_ops from the current coefficient fields
        # leaf adjacency on the Gauss nodes (the refinement residuals need it)
        self.nbr = torch.as_tensor(_neighbors(dom)).to(device)
        self.nb, self.nq = pde.P.shape
        self.n = dom.p**2 if dom.bool_2D else dom.p**3
        self.nl = dom.n_leaves
        self.nroot = root_points(dom)
        self.dtype = dtype
        self.device = device

    def check(self, u, f, g_root, g_leaf=None):
        """Shapes the residual kernels index with (they do no bounds checks)."""
        nl, n, w = u.shape
        ok = (nl, n) == (self.nl, self.n) and tuple(g_root.shape) == (self.nroot, w)
        ok = ok and (f is None or tuple(f.shape) == (nl, n, w))
        ok = ok and (g_leaf is None or tuple(g_leaf.shape) == (nl, self.nq, w))
        ok = ok and all(t is None or (t.dtype == self.dtype and t.is_contiguous()) for t in (u, f, g_root, g_leaf))
        if not ok:
            raise ValueError("refinement operands do not match the problem's shapes / dtype")

    def products(self, u):
        """R (nl, nrow_R, w) = [op_k u | extra] for the iterate u (nl, n, w)."""
        nl, n, w = u.shape
        if w == 1:
            return _lib.gemm(u.view(nl, n), self.MT).view(nl, self.nrow_R, 1)
        if w >= 16:
            return _lib.gemm(self.M, u)  # full 16-column product tiles
        # a few sources: one product over every (leaf, source) row rather
        # than nl products with w columns each (w of 16 tile columns used)
        ut = u.transpose(1, 2).reshape(nl * w, n).contiguous()
        return _lib.gemm(ut, self.MT).view(nl, w, self.nrow_R).transpose(1, 2).contiguous()


def root_points(dom) -> int:
    """Number of Gauss points on the root boundary of a uniform tree."""
    return 4 * dom.q * 2**dom.L if dom.bool_2D else 6 * dom.q**2 * 4**dom.L


@functools.lru_cache(maxsize=8)
def _neighbors_cached(bool_2D, L, q, bounds):
    if bool_2D:
        return gauss_neighbors_2D(L, q)
    from ._discretization_tree import DiscretizationNode3D

    return gauss_neighbors_3D(DiscretizationNode3D(*bounds), L, q)


def _neighbors(dom):
    r = dom.root
    bounds = (r.xmin, r.xmax, r.ymin, r.ymax) if dom.bool_2D else (r.xmin, r.xmax, r.ymin, r.ymax, r.zmin, r.zmax)
    return _neighbors_cached(dom.bool_2D, dom.L, dom.q, tuple(float(b) for b in bounds))


def tree_down(pde, g_root, g_lst, v, Y="Y"):
    """Down pass through the build's stored S (each level in its own
    precision) from root data (nroot, w); returns (u, leaf boundary data).
    3D roots are stored without their batch axis and use the oct tables."""
    from .down_pass import down_levels
    from ._tables import oct_merge_tables

    S_lst = list(pde.S_lst)
    g_lst = list(g_lst) if g_lst else []
    table_fn = None
    if not pde.domain.bool_2D:
        table_fn = lambda ne: oct_merge_tables(ne // 24)  # noqa: E731
        if S_lst[-1].ndim == 2:
            S_lst[-1] = S_lst[-1].unsqueeze(0)
            if g_lst and g_lst[-1].ndim == S_lst[-1].ndim - 1:
                g_lst[-1] = g_lst[-1].unsqueeze(0)
    return down_levels(g_root, S_lst, g_lst, getattr(pde, Y) if isinstance(Y, str) else Y, v, table_fn=table_fn)


def refine_ops(pde, device, dtype):
    """The operators are kept across PDEProblem.reset() (they depend on the
    domain and on which coefficient fields are present); the coefficient
    fields themselves are re-read after a reset."""
    names = tuple(nm for nm, _ in pde.coefficient_list())
    skey = ("refine_static", str(device), dtype, names)
    ops = pde._dev_cache.get(skey)
    if ops is None:
        ops = pde._dev_cache[skey] = RefineOps(pde, device, dtype)
    ckey = ("refine", str(device), dtype)
    if ckey not in pde._dev_cache:
        pde._dev_cache[ckey] = torch.stack([to_device(a, device, dtype) for _, a in pde.coefficient_list()]).contiguous()
    ops.coeffs = pde._dev_cache[ckey]
    return ops


def residual_iti(ops: RefineOps, u, f, g_root, out_dtype):
    """FP64 residual (nl, n, w) in out_dtype of the ItI system; returns (res, sum |res|^2 as a device scalar)."""
    lib = load_library()
    hd = _lib.handle()
    ops.check(u, f, g_root)
    nl, n, w = u.shape
    R = ops.products(u)
    res = torch.empty((nl, n, w), dtype=out_dtype, device=u.device)
    part = torch.empty((nl,), dtype=torch.float64, device=u.device)
    rc = lib.hps_residual_iti(
        hd, dtype_code(ops.dtype), dtype_code(out_dtype), nl, n, ops.nb, ops.nq, w, ops.nops, len(ops.coef_ops),
        ops.coef_ops.ctypes.data_as(ctypes.c_void_p), ptr(ops.coeffs), ptr(R), ops.nrow_R, ptr(u), ptr(f),
        ptr(ops.P), ptr(g_root), ptr(ops.nbr), ptr(res), ptr(part),
    )
    check(rc, hd, "hps_residual_iti")
    return res, part.sum()


def residual_dtn(ops: RefineOps, u, g_leaf, f, g_root, out_dtype):
    """FP64 DtN residuals: (res (nl, n, w), hinj (nl, nq, w), groot (nroot, w), sum of squares)."""
    lib = load_library()
    hd = _lib.handle()
    ops.check(u, f, g_root, g_leaf)
    nl, n, w = u.shape
    R = ops.products(u)
    res = torch.empty((nl, n, w), dtype=out_dtype, device=u.device)
    hinj = torch.empty((nl, ops.nq, w), dtype=out_dtype, device=u.device)
    groot = torch.empty((g_root.shape[0], w), dtype=out_dtype, device=u.device)
    part = torch.empty((nl,), dtype=torch.float64, device=u.device)
    rc = lib.hps_residual_dtn(
        hd, dtype_code(ops.dtype), dtype_code(out_dtype), nl, n, ops.nb, ops.nq, w, ops.nops, len(ops.coef_ops),
        ops.coef_ops.ctypes.data_as(ctypes.c_void_p), ptr(ops.coeffs), ptr(R), ops.nrow_R, ptr(u), ptr(f),
        ptr(g_leaf), ptr(g_root), ptr(ops.nbr), ptr(res), ptr(hinj), ptr(groot), ptr(part),
    )
    check(rc, hd, "hps_residual_dtn")
    return res, hinj, groot, part.sum()


def refine_update(x, dx):
    """x += dx in x's precision; returns (sum |dx|^2, sum |x|^2) as device scalars."""
    lib = load_library()
    hd = _lib.handle()
    part = torch.empty((2 * _NPARTS,), dtype=torch.float64, device=x.device)
    rc = lib.hps_refine_update(hd, dtype_code(x.dtype), dtype_code(dx.dtype), x.numel(), ptr(x), ptr(dx.contiguous()),
                               _NPARTS, ptr(part))
    check(rc, hd, "hps_refine_update")
    p = part.view(_NPARTS, 2).sum(0)
    return p[0], p[1]


def set_boundary_rows(ops: RefineOps, u, g_leaf):
    """u[:, :nb, :] = P g_leaf in u's precision (DtN iterate consistency)."""
    lib = load_library()
    hd = _lib.handle()
    nl, n, w = u.shape
    rc = lib.hps_gemm_batched(
        hd, dtype_code(u.dtype), ops.nb, w, ops.nq, 1.0, ptr(ops.P), ops.nq, 0, ptr(g_leaf), w, ops.nq * w, 0.0,
        None, w, 0, ptr(u), w, n * w, nl,
    )
    check(rc, hd, "hps_gemm_batched")


def _as3(x):
    return x if x.ndim == 3 else x.unsqueeze(-1)


def _root_data(bd, device, dtype, w):
    if isinstance(bd, (list, tuple)):
        bd = torch.cat([to_device(b, device, dtype) for b in bd])
    t = to_device(bd, device, dtype)
    if t.ndim == 1:
        t = t.unsqueeze(-1)
    if t.shape[-1] != w:
        t = t.expand(-1, w)
    return t.contiguous()


def refine_solve(pde, boundary_data, u0, source, g_leaf0=None, tol=1e-11, max_iter=10):
    """Refine the mixed-precision solution u0 (nl, n[, w]) to FP64.

    source: the FP64 source the solution is for (None = zero).  g_leaf0: the
    DtN leaves' Gauss data of the first solve.  Returns u in FP64 (c128 for
    complex problems) shaped like u0, and records the iteration history in
    ``pde.refine_history``.

    Stopping: after each correction the residual of the new iterate is
    evaluated; the next update is estimated as the last update times the
    residual's reduction and the loop ends when that estimate (or an actual
    update) is below tol relative to |u|."""
    from .up_pass import _up

    squeeze = u0.ndim == 2
    u0 = _as3(u0)
    dev = u0.device
    hi = HIGH[u0.dtype]
    lo = LOW[hi]
    ops = refine_ops(pde, dev, hi)
    nl, n, w = u0.shape
    u = cast(u0, hi).contiguous()
    f = None
    if source is not None:
        f = _as3(to_device(source, dev, hi))
        if f.shape[-1] != w:
            f = f.expand(-1, -1, w)
        f = f.contiguous()
    g_root = _root_data(boundary_data, dev, hi, w)
    nroot = root_points(pde.domain)
    if g_root.shape[0] != nroot:
        raise ValueError(f"boundary data has {g_root.shape[0]} points, the root boundary has {nroot}")
    iti = pde.use_ItI
    g_leaf = None
    if not iti:
        g_leaf = cast(_as3(g_leaf0), hi).contiguous()
        set_boundary_rows(ops, u, g_leaf)
    hist = []
    converged = False
    prev = None  # (residual norm, update norm) of the last correction
    for it in range(max_iter + 1):
        with profiling.stage("refine_residual"):
            if iti:
                res, rn2 = residual_iti(ops, u, f, g_root, lo)
            else:
                res, hinj, groot, rn2 = residual_dtn(ops, u, g_leaf, f, g_root, lo)
        if prev is not None:
            # The correction operator is linear, so the next update would be
            # about the last one scaled by the residual's reduction; stop
            # without computing it once that estimate is below tol.
            rn = float(rn2.sqrt())
            est = prev[1] * rn / max(prev[0], 1e-300)
            hist[-1]["residual_after"] = rn
            hist[-1]["next_update_estimate"] = est / max(un, 1e-300)
            if est <= tol * un:
                converged = True
                break
        if it == max_iter:
            break
        if iti:
            with profiling.stage("refine_up"):
                v, g_lst = _up(res, pde, rhs_full=True)
            with profiling.stage("refine_down"):
                du, _ = tree_down(pde, torch.zeros_like(g_root), g_lst, v)
            with profiling.stage("refine_update"):
                dn2, un2 = refine_update(u, du)
        else:
            with profiling.stage("refine_up"):
                v, g_lst = _up(res, pde, rhs_full=True, hinj=hinj)
            with profiling.stage("refine_down"):
                du, dg = tree_down(pde, groot, g_lst, v)
            with profiling.stage("refine_update"):
                dn2, un2 = refine_update(u, du)
                refine_update(g_leaf, dg)
                set_boundary_rows(ops, u, g_leaf)
        dn, un, rn = (float(x) for x in torch.stack([dn2, un2, rn2]).sqrt().cpu())
        hist.append({"iter": it + 1, "residual": rn, "rel_update": dn / max(un, 1e-300)})
        prev = (rn, dn)
        if dn <= tol * un:
            converged = True
            break
    pde.refine_history = hist
    pde.refine_converged = converged
    if not converged:
        warnings.warn(f"iterative refinement stopped after {max_iter} iterations at relative update "
                      f"{hist[-1]['rel_update']:.2e} (tol {tol:.1e})", RuntimeWarning, stacklevel=3)
    return u[..., 0] if squeeze else u
