"""HPS grid <-> rectangular grid: ``Domain.interp_to_interior_points`` and
``Domain.interp_from_interior_points`` (``_domain.py:99-294``,
``_interpolation_methods.py:24-340``).

To the HPS grid: the reference builds one dense (p^d x n_x n_y [n_z])
barycentric matrix per leaf.  It is a tensor product, so here every leaf
column's x factor (p x n_x) is stacked into one matrix and the whole grid of
leaf values comes from two (three in 3D) GEMMs on the device; each leaf then
gathers its p^d block in its point order.

From the HPS grid: the HIP kernel ``hps_interp_grid`` takes the grid's axes,
finds each point's leaf -- the first in leaf order whose closed box holds it
(the reference's argmax over containment) -- and evaluates the point's
barycentric interpolant on that leaf.
"""

from __future__ import annotations

import numpy as np
import torch

from . import _lib
from ._discretization_tree import get_all_leaves
from ._grid_creation import _subdivide_2D, _subdivide_3D, rearrange_indices_ext_int_2D, rearrange_indices_ext_int_3D
from ._lib import HPS_C128, HPS_F64, check, load_library, ptr
from .quadrature import _bary_factor_eps, affine_transform, chebyshev_points


def _device(device):
    return torch.device(device) if device is not None else torch.device("cuda", 0)


def _factors(samples, intervals, p, flip):
    """Stacked 1D barycentric factors (n_side p, n_samples): block i maps the
    samples to cell column i's Chebyshev nodes (reversed when flip)."""
    c = chebyshev_points(p)
    blocks = []
    for lo_hi in intervals:
        nodes = affine_transform(c, np.asarray(lo_hi))
        if flip:
            nodes = np.flip(nodes)
        blocks.append(_bary_factor_eps(samples, nodes))
    return np.concatenate(blocks, axis=0)


def _leaf_boxes(domain):
    """(n_leaves, 2 d) [lo, hi] per coordinate of the leaves, in leaf order."""
    cache = domain.__dict__.setdefault("_interp_cache", {})
    if "boxes" not in cache:
        r = domain.root
        if not domain.bool_uniform:
            keys = ("xmin", "xmax", "ymin", "ymax") + (() if domain.bool_2D else ("zmin", "zmax"))
            b = [[getattr(lf, k) for k in keys] for lf in get_all_leaves(r)]
        elif domain.bool_2D:
            b = _subdivide_2D([r.xmin, r.xmax, r.ymin, r.ymax], domain.L)
        else:
            b = _subdivide_3D([r.xmin, r.xmax, r.ymin, r.ymax, r.zmin, r.zmax], domain.L)
        cache["boxes"] = np.asarray(b, dtype=np.float64).reshape(-1, 4 if domain.bool_2D else 6)
    return cache["boxes"]


def _axis_intervals(boxes, axis):
    """The distinct leaf intervals [lo, hi] along axis (sorted) and each
    leaf's index into them."""
    uniq, inv = np.unique(boxes[:, 2 * axis:2 * axis + 2], axis=0, return_inverse=True)
    return uniq, np.asarray(inv).reshape(-1)


def _leaf_gather(domain, boxes, dev):
    """(n_leaves, p^d) flat indices into the tensor grid of every distinct
    interval's points (x slowest) in leaf point order; cached on the domain."""
    cache = domain.__dict__.setdefault("_interp_cache", {})
    key = ("gather", str(dev))
    if key not in cache:
        p = domain.p
        d = boxes.shape[1] // 2
        ivs = [_axis_intervals(boxes, a) for a in range(d)]
        sizes = [len(u) * p for u, _ in ivs]
        if d == 2:
            a, b = np.divmod(rearrange_indices_ext_int_2D(p), p)
            idx = (ivs[0][1][:, None] * p + a[None, :]) * sizes[1] + (ivs[1][1][:, None] * p + b[None, :])
        else:
            a, rem = np.divmod(rearrange_indices_ext_int_3D(p), p * p)
            b, c = np.divmod(rem, p)
            idx = ((ivs[0][1][:, None] * p + a[None, :]) * sizes[1] + (ivs[1][1][:, None] * p + b[None, :])) * sizes[
                2] + (ivs[2][1][:, None] * p + c[None, :])
        cache[key] = torch.as_tensor(idx, device=dev)
    return cache[key]


def interp_to_interior_points(domain, values, sample_points_x, sample_points_y, sample_points_z=None, device=None):
    """Values on meshgrid(x, y[, z], indexing="ij") -> (n_leaves, p^d) on the device."""
    dev = _device(device)
    x = np.asarray(sample_points_x, dtype=np.float64)
    y = np.asarray(sample_points_y, dtype=np.float64)
    v = torch.as_tensor(values)
    dt = torch.complex128 if v.is_complex() else torch.float64
    v = v.to(device=dev, dtype=dt).contiguous()
    p = domain.p
    boxes = _leaf_boxes(domain)
    if domain.bool_2D:
        if sample_points_z is not None:
            raise ValueError("2D domain: sample_points_z must be None")
        if tuple(v.shape) != (x.size, y.size):
            raise ValueError(f"values must have shape {(x.size, y.size)}")
        Fx = torch.as_tensor(_factors(x, _axis_intervals(boxes, 0)[0], p, False)).to(dev, dt).contiguous()
        Fy = torch.as_tensor(_factors(y, _axis_intervals(boxes, 1)[0], p, True)).to(dev, dt)
        T1 = _lib.gemm(Fx, v)[0]  # (n_xint p, n_y)
        G = _lib.gemm(T1, Fy.t().contiguous())[0]  # (n_xint p, n_yint p)
        return G.reshape(-1)[_leaf_gather(domain, boxes, dev)]
    z = np.asarray(sample_points_z, dtype=np.float64)
    if tuple(v.shape) != (x.size, y.size, z.size):
        raise ValueError(f"values must have shape {(x.size, y.size, z.size)}")
    Fx = torch.as_tensor(_factors(x, _axis_intervals(boxes, 0)[0], p, False)).to(dev, dt).contiguous()
    Fy = torch.as_tensor(_factors(y, _axis_intervals(boxes, 1)[0], p, False)).to(dev, dt).contiguous()
    Fz = torch.as_tensor(_factors(z, _axis_intervals(boxes, 2)[0], p, False)).to(dev, dt)
    Px, Py, Pz = Fx.shape[0], Fy.shape[0], Fz.shape[0]
    T1 = _lib.gemm(Fx, v.reshape(x.size, -1))[0].reshape(Px, y.size, z.size)  # x contracted
    T2 = _lib.gemm(Fy, T1.permute(1, 0, 2).reshape(y.size, -1).contiguous())[0]  # (Py, Px n_z)
    T2 = T2.reshape(Py, Px, z.size).permute(1, 0, 2).reshape(Px * Py, z.size).contiguous()  # (Px Py, n_z)
    G = _lib.gemm(T2, Fz.t().contiguous())[0]  # (Px Py, Pz)
    return G.reshape(-1)[_leaf_gather(domain, boxes, dev)]


def _grid_consts(domain, dev):
    """Device tables of a domain for hps_interp_grid, cached on it: the leaf
    of every cell of the uniform grid at the finest leaf size."""
    cache = domain.__dict__.setdefault("_interp_cache", {})
    key = str(dev)
    if key not in cache:
        p = domain.p
        boxes = _leaf_boxes(domain)
        d = boxes.shape[1] // 2
        lo = boxes[:, 0::2].min(axis=0)
        hi = boxes[:, 1::2].max(axis=0)
        width = (boxes[:, 1] - boxes[:, 0]).min()
        n_side = int(round((hi[0] - lo[0]) / width))
        h = (hi - lo) / n_side
        leaf_of_cell = np.empty((n_side,) * d, dtype=np.int32)
        c0 = np.rint((boxes[:, 0::2] - lo) / h).astype(np.int64)
        c1 = np.rint((boxes[:, 1::2] - lo) / h).astype(np.int64)
        for l in range(boxes.shape[0]):
            leaf_of_cell[tuple(slice(c0[l, a], c1[l, a]) for a in range(d))] = l
        root = np.stack([lo, hi], axis=1).reshape(-1)
        r = rearrange_indices_ext_int_2D(p) if d == 2 else rearrange_indices_ext_int_3D(p)
        grid_to_leaf = np.empty(r.size, dtype=np.int32)
        grid_to_leaf[r] = np.arange(r.size, dtype=np.int32)

        def put(a, t):
            return torch.as_tensor(np.ascontiguousarray(a)).to(device=dev, dtype=t)

        cache[key] = dict(n_side=n_side, d=d, boxes=put(boxes, torch.float64), root=put(root, torch.float64),
                          leaf_of_cell=put(leaf_of_cell, torch.int32), cheb=put(chebyshev_points(p), torch.float64),
                          grid_to_leaf=put(grid_to_leaf, torch.int32))
    return cache[key]


def interp_from_interior_points(domain, samples, eval_points_x, eval_points_y, eval_points_z=None, device=None):
    """(vals, target_pts) as the reference returns them: vals (n_x, n_y[, nsrc])
    (3D: (n_x, n_y, n_z)) from the points of meshgrid(x, y[, z]) in its default
    ("xy") order, and target_pts ((n_y, n_x, 2) or (n_y, n_x, n_z, 3)).  The
    grid points, their leaves and the interpolants are all computed on the
    device (``hps_interp_grid``)."""
    dev = _device(device)
    s = torch.as_tensor(samples)
    dt = torch.complex128 if s.is_complex() else torch.float64
    s = s.to(device=dev, dtype=dt)
    nl, npl = domain.interior_points.shape[:2]
    if tuple(s.shape[:2]) != (nl, npl):
        raise ValueError(f"samples must have leading shape {(nl, npl)}")
    multi = s.dim() == 3
    s3 = (s if multi else s.unsqueeze(-1)).contiguous()
    nsrc = s3.shape[-1]
    x = np.asarray(eval_points_x, dtype=np.float64).reshape(-1)
    y = np.asarray(eval_points_y, dtype=np.float64).reshape(-1)
    c = _grid_consts(domain, dev)
    if domain.bool_2D:
        if eval_points_z is not None:
            raise ValueError("2D domain: eval_points_z must be None")
        z = np.zeros(1)
        X, Y = np.meshgrid(x, y)
        target = np.stack([X, Y], axis=2)
        out_shape = (x.size, y.size)
    else:
        z = np.asarray(eval_points_z, dtype=np.float64).reshape(-1)
        X, Y, Z = np.meshgrid(x, y, z)
        target = np.concatenate([X[..., None], Y[..., None], Z[..., None]], axis=3)
        out_shape = (x.size, y.size, z.size)
    nz = 1 if domain.bool_2D else z.size
    npts = x.size * y.size * nz
    xs, ys, zs = (torch.as_tensor(a).to(dev) for a in (x, y, z))
    out = torch.empty((npts, nsrc), dtype=dt, device=dev)
    lib = load_library()
    hd = _lib.handle()
    rc = lib.hps_interp_grid(hd, HPS_C128 if dt == torch.complex128 else HPS_F64, x.size, y.size, nz, c["d"],
                             domain.p, nsrc, ptr(xs), ptr(ys), ptr(zs), c["n_side"], ptr(c["root"]),
                             ptr(c["leaf_of_cell"]), ptr(c["boxes"]), ptr(c["cheb"]), 1, ptr(c["grid_to_leaf"]),
                             ptr(s3), ptr(out))
    check(rc, hd, "hps_interp_grid")
    vals = out.reshape(out_shape + (nsrc,)) if multi else out[:, 0].reshape(out_shape)
    return vals, target


__all__ = ["interp_to_interior_points", "interp_from_interior_points"]
