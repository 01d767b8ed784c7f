"""Adaptive (non-uniform) 2D quadtrees: refinement, grids and the DtN
build / solve on them.

Follows the reference's adaptive 2D path: ``_adaptive_discretization_2D.py``
(refinement with level restriction), ``_discretization_tree_operations_2D.py``
(node splitting and boundary counts), ``_grid_creation_2D.py:128-201``
(points), ``local_solve/_adaptive_2D_DtN.py``, ``merge/_adaptive_2D_DtN.py``
with ``merge/_utils_adaptive_2D_DtN.py``, and ``down_pass/_adaptive_2D_DtN.py``.

Design on the device:

* Leaves of every size go through ONE batched ``hps_leaf_solve`` with the
  operators of half side 1: a k-th order coefficient of a leaf of half side
  s is scaled by s^-k and its T and h by 1/s (``hps_scale_batched``), which
  is the reference's per-leaf operator rescaling folded into the data.
* Where two children's discretisations differ along an interface, the finer
  child's map is compressed, T' = R T P and h' = R h, with R (P) the block
  diagonal of identities and the two-panels-to-one (one-to-two) Gauss
  interpolations L_1f2 (L_2f1) on its flagged panel pairs.  The compressed
  children are packed into padded blocks and the 4:1 merge is the ordinary
  index-table merge ``hps_merge_level``; merges with the same children sizes
  and compression pattern at a level share their tables and run as one batch.
* The down pass scatters with ``hps_down_level`` through the same tables and
  refines compressed interface data with P.
"""

from __future__ import annotations

import ctypes
from collections import OrderedDict

import numpy as np
import torch

from . import _lib
from ._discretization_tree import DiscretizationNode2D, get_all_leaves, get_depth, get_nodes_at_level
from ._grid_creation import bounds_to_cheby_points_2D, rearrange_indices_ext_int_2D
from ._lib import HPS_KIND_DTN, check, dtype_code, load_library, ptr
from .quadrature import (
    affine_transform,
    barycentric_lagrange_interpolation_matrix_1D,
    chebyshev_points,
    chebyshev_weights,
    gauss_points,
)

_A, _B, _C, _D = 0, 1, 2, 3


# ---------------------------------------------------------------------------
# tree operations (_discretization_tree_operations_2D.py)
# ---------------------------------------------------------------------------
def get_four_children(parent):
    """Children SW, SE, NE, NW one level deeper."""
    xm = (parent.xmin + parent.xmax) / 2
    ym = (parent.ymin + parent.ymax) / 2
    d = parent.depth + 1
    return (
        DiscretizationNode2D(xmin=parent.xmin, xmax=xm, ymin=parent.ymin, ymax=ym, depth=d),
        DiscretizationNode2D(xmin=xm, xmax=parent.xmax, ymin=parent.ymin, ymax=ym, depth=d),
        DiscretizationNode2D(xmin=xm, xmax=parent.xmax, ymin=ym, ymax=parent.ymax, depth=d),
        DiscretizationNode2D(xmin=parent.xmin, xmax=xm, ymin=ym, ymax=parent.ymax, depth=d),
    )


def _path_to(root, node):
    """Ancestors of node from the root down to its parent."""
    path = []
    cur = root
    cx, cy = (node.xmin + node.xmax) / 2, (node.ymin + node.ymax) / 2
    while cur is not node:
        if not cur.children:
            raise ValueError("node is not in the tree under root")
        path.append(cur)
        xm, ym = (cur.xmin + cur.xmax) / 2, (cur.ymin + cur.ymax) / 2
        cur = cur.children[(0 if cy < ym else 3) if cx < xm else (1 if cy < ym else 2)]
    return path


def add_four_children(add_to, root=None, q: int | None = None) -> None:
    """Split a leaf into four children.  With root and q, keep the boundary
    counts n_0..n_3 (Gauss points on sides S, E, N, W) of add_to and of its
    ancestors up to date (``_discretization_tree_operations_2D.py:50-107``)."""
    if len(add_to.children):
        return
    add_to.children = get_four_children(add_to)
    if q is None:
        return
    for ch in add_to.children:
        ch.n_0 = ch.n_1 = ch.n_2 = ch.n_3 = q
    add_to.n_0 = add_to.n_1 = add_to.n_2 = add_to.n_3 = 2 * q
    if root is None or root is add_to:
        return
    for nd in reversed(_path_to(root, add_to)):
        a, b, c, d = nd.children
        nd.n_0, nd.n_1, nd.n_2, nd.n_3 = a.n_0 + b.n_0, b.n_1 + c.n_1, c.n_2 + d.n_2, d.n_3 + a.n_3


def _side_counts(node, q):
    """(n_0, n_1, n_2, n_3) = q x the leaves on each side, from the geometry."""
    if not len(node.children):
        return (q, q, q, q)
    a, b, c, d = (_side_counts(ch, q) for ch in node.children)
    return (a[0] + b[0], b[1] + c[1], c[2] + d[2], d[3] + a[3])


def assign_side_counts(root, q: int) -> None:
    """Set n_0..n_3 on every node of the tree."""
    def rec(nd):
        if not len(nd.children):
            nd.n_0 = nd.n_1 = nd.n_2 = nd.n_3 = q
            return
        for ch in nd.children:
            rec(ch)
        a, b, c, d = nd.children
        nd.n_0, nd.n_1, nd.n_2, nd.n_3 = a.n_0 + b.n_0, b.n_1 + c.n_1, c.n_2 + d.n_2, d.n_3 + a.n_3

    rec(root)


# ---------------------------------------------------------------------------
# grids (_grid_creation_2D.py:128-201)
# ---------------------------------------------------------------------------
def _bounds(node):
    return np.array([node.xmin, node.xmax, node.ymin, node.ymax], dtype=np.float64)


def compute_interior_Chebyshev_points_adaptive_2D(root, p: int) -> np.ndarray:
    c = chebyshev_points(p)
    return np.stack([bounds_to_cheby_points_2D(_bounds(lf), c) for lf in get_all_leaves(root)])


def get_ordered_lst_of_boundary_nodes(root):
    """Leaves along the root's sides S (W->E), E (S->N), N (E->W), W (N->S)."""
    lv = get_all_leaves(root)
    s = sorted([x for x in lv if x.ymin == root.ymin], key=lambda x: x.xmin)
    e = sorted([x for x in lv if x.xmax == root.xmax], key=lambda x: x.ymin)
    n = sorted([x for x in lv if x.ymax == root.ymax], key=lambda x: -x.xmin)
    w = sorted([x for x in lv if x.xmin == root.xmin], key=lambda x: -x.ymin)
    return s, e, n, w


def compute_boundary_Gauss_points_adaptive_2D(root, q: int) -> np.ndarray:
    g = gauss_points(q)
    s, e, n, w = get_ordered_lst_of_boundary_nodes(root)
    xs = np.concatenate([affine_transform(g, [x.xmin, x.xmax]) for x in s])
    ye = np.concatenate([affine_transform(g, [x.ymin, x.ymax]) for x in e])
    xn = np.concatenate([affine_transform(g, [x.xmax, x.xmin]) for x in n])
    yw = np.concatenate([affine_transform(g, [x.ymax, x.ymin]) for x in w])
    return np.concatenate([
        np.stack([xs, np.full_like(xs, root.ymin)], -1),
        np.stack([np.full_like(ye, root.xmax), ye], -1),
        np.stack([xn, np.full_like(xn, root.ymax)], -1),
        np.stack([np.full_like(yw, root.xmin), yw], -1),
    ])


# ---------------------------------------------------------------------------
# adaptive refinement (_adaptive_discretization_2D.py)
# ---------------------------------------------------------------------------
def node_to_bounds(node) -> np.ndarray:
    return _bounds(node)


def precompute_L_4f1(p: int) -> np.ndarray:
    """(4 p^2, p^2) interpolation from a leaf's Chebyshev grid to its four
    children's grids (SW, SE, NE, NW), every grid in leaf point order."""
    c = chebyshev_points(p)
    r = rearrange_indices_ext_int_2D(p)
    i, j = np.divmod(r, p)
    ix, iy = i, p - 1 - j  # 1D x / y node of each leaf point
    X, Y = c[ix], c[iy]
    blocks = []
    for cx, cy in ((-0.5, -0.5), (0.5, -0.5), (0.5, 0.5), (-0.5, 0.5)):
        Lx = barycentric_lagrange_interpolation_matrix_1D(c, cx + 0.5 * X)
        Ly = barycentric_lagrange_interpolation_matrix_1D(c, cy + 0.5 * Y)
        blocks.append(Lx[:, ix] * Ly[:, iy])
    return np.concatenate(blocks, axis=0)


def get_squared_l2_norm_single_panel(f_evals, bounds, p: int) -> float:
    """Clenshaw-Curtis estimate of the squared L2 norm of f on one panel
    from its leaf-ordered samples (``_adaptive_discretization_2D.py:384-407``)."""
    b = np.asarray(bounds, dtype=np.float64)
    w = np.outer(chebyshev_weights(p, b[:2]), chebyshev_weights(p, b[2:])).reshape(-1)
    w = w[rearrange_indices_ext_int_2D(p)]
    return float(np.sum(np.asarray(f_evals) ** 2 * w))


def _panel_points(B, p):
    """Leaf points of the panels B (m, 4) -> (m, p^2, 2)."""
    c = chebyshev_points(p)
    return np.stack([bounds_to_cheby_points_2D(b, c) for b in B]) if len(B) else np.zeros((0, p * p, 2))


def _child_bounds(B):
    xm, ym = (B[:, 0] + B[:, 1]) / 2, (B[:, 2] + B[:, 3]) / 2
    kids = [
        np.stack([B[:, 0], xm, B[:, 2], ym], -1),
        np.stack([xm, B[:, 1], B[:, 2], ym], -1),
        np.stack([xm, B[:, 1], ym, B[:, 3]], -1),
        np.stack([B[:, 0], xm, ym, B[:, 3]], -1),
    ]
    return np.stack(kids, axis=1)  # (m, 4, 4)


def patches_to_check(newly_refined, root):
    """Boxes of newly_refined's size adjacent to it (W, E, S, N) inside the
    root (``_adaptive_discretization_2D.py:212-274``)."""
    s = newly_refined.xmax - newly_refined.xmin
    n = newly_refined
    out = []
    if n.xmin - s >= root.xmin:
        out.append((n.xmin - s, n.xmin, n.ymin, n.ymax))
    if n.xmax + s <= root.xmax:
        out.append((n.xmax, n.xmax + s, n.ymin, n.ymax))
    if n.ymin - s >= root.ymin:
        out.append((n.xmin, n.xmax, n.ymin - s, n.ymin))
    if n.ymax + s <= root.ymax:
        out.append((n.xmin, n.xmax, n.ymax, n.ymax + s))
    return out


def find_or_add_child_2D(node, root, q, xmin, xmax, ymin, ymax):
    """Make the quadtree cell [xmin, xmax] x [ymin, ymax] exist below node by
    splitting the leaf that covers it; returns the split node or None
    (``_adaptive_discretization_2D.py:277-359``)."""
    want = xmax - xmin
    side = node.xmax - node.xmin
    if np.isclose(2 * want, side, rtol=1e-5, atol=1e-8):
        if not len(node.children):
            add_four_children(node, root=root, q=q)
            return node
        return None
    xm, ym = (node.xmin + node.xmax) / 2, (node.ymin + node.ymax) / 2
    child = node.children[(0 if ymin < ym else 3) if xmin < xm else (1 if ymin < ym else 2)]
    if np.isclose(side, 4 * want, rtol=1e-5, atol=1e-8):
        if not len(child.children):
            add_four_children(child, root=root, q=q)
            return child
        return None
    if side < 2 * want:
        raise ValueError("Requested volume is too large for the current node")
    return find_or_add_child_2D(child, root, q, xmin, xmax, ymin, ymax)


def generate_adaptive_mesh_level_restriction_2D(root, f_fn, tol: float, p: int, q: int, restrict_bool: bool = True,
                                                l2_norm: bool = False) -> None:
    """Refine leaves until f is resolved: a leaf passes when interpolating its
    samples to its children's grids matches f there to tol, relative to the
    running max-norm of f (or, with l2_norm, in squared L2 relative to f's
    squared L2 norm).  With restrict_bool, every split also splits the
    same-size neighbours' parents as needed so that adjacent leaves differ
    by at most one level."""
    R = precompute_L_4f1(p)
    m = p * p
    if l2_norm:
        add_four_children(root, root=root, q=q)
        lv = get_all_leaves(root)
        fv = np.asarray(f_fn(_panel_points([_bounds(x) for x in lv], p)))
        gnorm = sum(get_squared_l2_norm_single_panel(fv[i], _bounds(x), p) for i, x in enumerate(lv))
        tol = tol**2
    else:
        kids = _child_bounds(_bounds(root)[None])[0]
        gnorm = float(np.max(np.asarray(f_fn(_panel_points(kids, p)))))
    queue = list(get_all_leaves(root))
    while queue:
        B = np.stack([_bounds(x) for x in queue])
        K = _child_bounds(B)
        f0 = np.asarray(f_fn(_panel_points(B, p)))  # (nq, m)
        f1 = np.asarray(f_fn(_panel_points(K.reshape(-1, 4), p))).reshape(len(queue), 4 * m)
        diff = f0 @ R.T - f1
        if l2_norm:
            err = np.array([sum(get_squared_l2_norm_single_panel(diff[i, k * m:(k + 1) * m], K[i, k], p)
                                for k in range(4)) for i in range(len(queue))])
            ok = err / gnorm < tol
        else:
            c = np.maximum(gnorm, np.max(np.abs(f1), axis=1))
            ok = np.max(np.abs(diff), axis=1) / c < tol
            gnorm = float(np.max(c))
        nxt = []
        for nd, good in zip(queue, ok):
            if good:
                continue
            add_four_children(nd, root=root, q=q)
            nxt.extend(nd.children)
            if restrict_bool:
                check_q = [nd]
                while check_q:
                    cur = check_q.pop()
                    for box in patches_to_check(cur, root):
                        made = find_or_add_child_2D(root, root, q, *box)
                        if made is not None:
                            check_q.append(made)
                            nxt.extend(made.children)
        queue = nxt


# ---------------------------------------------------------------------------
# merge plans: compression operators and index tables
# ---------------------------------------------------------------------------
def precompute_projection_ops_2D(q: int):
    """L_2f1 (2q x q): one Gauss panel to two; L_1f2 (q x 2q): two to one
    (``_precompute_operators_2D.py:269-300``)."""
    g = gauss_points(q)
    g2 = np.concatenate([affine_transform(g, [-1.0, 0.0]), affine_transform(g, [0.0, 1.0])])
    return barycentric_lagrange_interpolation_matrix_1D(g, g2), barycentric_lagrange_interpolation_matrix_1D(g2, g)


def _interface_flags(first, second, inner_end, sign):
    """Panel compression flags of two leaf lists along one interface, both
    ordered from the outer boundary inwards (``merge/_utils_adaptive_2D_DtN.py:
    10-250``): a pair of leaves on one side facing one on the other gets
    compressed."""
    f1, f2 = [], []
    i = j = 0
    while i < len(first) and j < len(second):
        e1, e2 = inner_end(first[i]), inner_end(second[j])
        if e1 == e2:
            f1.append(False)
            f2.append(False)
            i += 1
            j += 1
        elif (e1 - e2) * sign > 0:
            f1 += [True, True]
            f2.append(False)
            i += 2
            j += 1
        else:
            f2 += [True, True]
            f1.append(False)
            j += 2
            i += 1
    return f1, f2


def find_compression_lists_2D(node_a, node_b, node_c, node_d):
    """Flags for interfaces a5, b5, b6, c6, c7, d7, d8, a8 (outside -> inside)."""
    la, lb, lc, ld = (get_all_leaves(x) for x in (node_a, node_b, node_c, node_d))
    a5 = sorted([x for x in la if x.xmax == node_a.xmax], key=lambda x: x.ymin)
    b5 = sorted([x for x in lb if x.xmin == node_b.xmin], key=lambda x: x.ymin)
    b6 = sorted([x for x in lb if x.ymax == node_b.ymax], key=lambda x: -x.xmin)
    c6 = sorted([x for x in lc if x.ymin == node_c.ymin], key=lambda x: -x.xmin)
    c7 = sorted([x for x in lc if x.xmin == node_c.xmin], key=lambda x: -x.ymin)
    d7 = sorted([x for x in ld if x.xmax == node_d.xmax], key=lambda x: -x.ymin)
    d8 = sorted([x for x in ld if x.ymin == node_d.ymin], key=lambda x: x.xmin)
    a8 = sorted([x for x in la if x.ymax == node_a.ymax], key=lambda x: x.xmin)
    fa5, fb5 = _interface_flags(a5, b5, lambda x: x.ymax, -1)
    fb6, fc6 = _interface_flags(b6, c6, lambda x: x.xmin, +1)
    fc7, fd7 = _interface_flags(c7, d7, lambda x: x.ymin, +1)
    fd8, fa8 = _interface_flags(d8, a8, lambda x: x.xmax, -1)
    return fa5, fb5, fb6, fc6, fc7, fd7, fd8, fa8


# side roles per child (sides S, E, N, W): None = exterior, else
# (interface, index of its flags in the compression lists, reversed): a side
# is reversed when its boundary order runs from the centre outwards
_ROLES = (
    (None, (5, 0, False), (8, 7, True), None),
    (None, None, (6, 2, False), (5, 1, True)),
    ((6, 3, True), None, None, (7, 4, False)),
    ((8, 6, False), (7, 5, True), None, None),
)
# the two (child, side) pairs of interfaces 5..8
_IFACES = {5: ((_A, 1), (_B, 3)), 6: ((_B, 2), (_C, 0)), 7: ((_C, 3), (_D, 1)), 8: ((_D, 0), (_A, 2))}
# merged exterior in final order: a_S, b_S, b_E, c_E, c_N, d_N, d_W, a_W
_EXT = ((_A, 0), (_B, 0), (_B, 1), (_C, 1), (_C, 2), (_D, 2), (_D, 3), (_A, 3))


class _Plan:
    """Tables and compression operators shared by the merges of one
    signature (children's side counts and compression flags)."""

    def __init__(self, counts, flags, q, L21, L12):
        J = lambda k: np.eye(k)[::-1]
        self.n = [sum(c) for c in counts]
        self.R, self.P = [], []
        off, length = [], []
        for ch in range(4):
            n_c = self.n[ch]
            rows_R, blocks = 0, []
            so, sl = [], []
            col = 0
            for side in range(4):
                role = _ROLES[ch][side]
                npan = counts[ch][side] // q
                fl = [False] * npan if role is None else list(flags[role[1]])
                rev = role is not None and role[2]
                if rev:
                    fl = fl[::-1]
                so.append(rows_R)
                i = 0
                while i < npan:
                    if fl[i]:
                        r = J(q) @ L12 @ J(2 * q) if rev else L12
                        p_ = J(2 * q) @ L21 @ J(q) if rev else L21
                        blocks.append((rows_R, col, r, p_))
                        rows_R += q
                        col += 2 * q
                        i += 2
                    else:
                        blocks.append((rows_R, col, None, None))
                        rows_R += q
                        col += q
                        i += 1
                sl.append(rows_R - so[-1])
            assert col == n_c
            off.append(so)
            length.append(sl)
            if rows_R == n_c:
                self.R.append(None)
                self.P.append(None)
                continue
            R = np.zeros((rows_R, n_c))
            P = np.zeros((n_c, rows_R))
            for r0, c0, r, p_ in blocks:
                if r is None:
                    R[r0:r0 + q, c0:c0 + q] = np.eye(q)
                    P[c0:c0 + q, r0:r0 + q] = np.eye(q)
                else:
                    R[r0:r0 + q, c0:c0 + 2 * q] = r
                    P[c0:c0 + 2 * q, r0:r0 + q] = p_
            self.R.append(R)
            self.P.append(P)
        self.m = [length[ch][0] + length[ch][1] + length[ch][2] + length[ch][3] for ch in range(4)]
        self.nmax = max(self.m)

        def local(ch, side, t):
            L = length[ch][side]
            rev = _ROLES[ch][side] is not None and _ROLES[ch][side][2]
            return off[ch][side] + (L - 1 - t if rev else t)

        pairs = []
        for iface in (5, 6, 7, 8):
            (c1, s1), (c2, s2) = _IFACES[iface]
            L = length[c1][s1]
            assert L == length[c2][s2], "interface discretisations do not match"
            pairs += [((c1, local(c1, s1, t)), (c2, local(c2, s2, t))) for t in range(L)]
        ext = [(ch, off[ch][side] + t) for ch, side in _EXT for t in range(length[ch][side])]
        self.ni, self.ne = len(pairs), len(ext)
        self.int_rows = _pack(pairs)
        self.int_cols = self.int_rows.copy()
        self.ext = _pack([[e] for e in ext])
        src = np.zeros(4 * self.nmax, dtype=np.int32)
        for k, (ch, l) in enumerate(ext):
            src[ch * self.nmax + l] = k
        for j, pr in enumerate(pairs):
            for ch, l in pr:
                src[ch * self.nmax + l] = -(j + 1)
        self.child_src = src
        self._dev = {}

    # the attributes and device() view the merge / down-pass wrappers read
    kind = "dtn"
    nc = 4

    @property
    def nec(self):
        return self.nmax

    @property
    def cols_final(self):
        return self.int_cols

    @property
    def final_of_internal(self):
        return np.arange(self.ni, dtype=np.int32)

    upd = np.zeros((0, 4), np.int32)
    d12_blk = np.zeros((0, 4), np.int32)
    d21_blk = np.zeros((0, 4), np.int32)

    def device(self, dev):
        key = str(dev)
        if key not in self._dev:
            self._dev[key] = tuple(
                torch.from_numpy(np.ascontiguousarray(x, dtype=np.int32)).to(dev)
                for x in (self.int_rows, self.int_cols, self.ext, self.child_src, self.cols_final,
                          self.final_of_internal)
            )
        return self._dev[key]

    def ops(self, ch, dev, dtype):
        key = ("RP", ch, str(dev), dtype)
        if key not in self._dev:
            self._dev[key] = (torch.as_tensor(self.R[ch]).to(dev, dtype).contiguous(),
                              torch.as_tensor(self.P[ch]).to(dev, dtype).contiguous())
        return self._dev[key]


def _pack(pairs):
    out = np.full((len(pairs), 4), -1, dtype=np.int32)
    out[:, 1] = 0
    out[:, 3] = 0
    for r, pr in enumerate(pairs):
        for a, (c, l) in enumerate(pr):
            out[r, 2 * a] = c
            out[r, 2 * a + 1] = l
    return out


def _signature(node, q):
    counts = tuple(_side_counts(ch, q) for ch in node.children)
    flags = tuple(tuple(f) for f in find_compression_lists_2D(*node.children))
    return counts, flags


# ---------------------------------------------------------------------------
# device helpers
# ---------------------------------------------------------------------------
def _gemm_strided(M, N, K, A, lda, sA, B, ldb, sB, D_ptr, ldd, sD, batch, dt):
    """D[b] = A[b] @ B[b] with explicit strides (sA = 0 broadcasts A)."""
    lib = load_library()
    hd = _lib.handle()
    rc = lib.hps_gemm_batched(hd, dtype_code(dt), M, N, K, 1.0, A, lda, sA, B, ldb, sB, 0.0, None, ldd, 0, D_ptr,
                              ldd, sD, batch)
    check(rc, hd, "hps_gemm_batched")


class _Packer:
    """Collects block copies (device addresses) and issues them as one
    hps_pack_blocks launch."""

    def __init__(self, dt, dev):
        self.dt, self.dev, self.rows, self.keep = dt, dev, [], []

    def add(self, src, dst_addr, rows, cols, lds, ldd):
        """Copy the rows x cols block of src (row stride lds) to dst_addr
        (row stride ldd); src must hold that block from its data pointer."""
        if src.dtype != self.dt or src.device != self.dev:
            raise ValueError("pack source has the wrong dtype or device")
        if src.numel() < (rows - 1) * lds + cols:
            raise ValueError("pack source is smaller than the block")
        self.rows.append((src.data_ptr(), dst_addr, rows, cols, lds, ldd))
        self.keep.append(src)

    def run(self):
        if not self.rows:
            return
        desc = torch.as_tensor(np.array(self.rows, dtype=np.int64)).to(self.dev)
        lib = load_library()
        hd = _lib.handle()
        rc = lib.hps_pack_blocks(hd, dtype_code(self.dt), len(self.rows), max(r[2] * r[3] for r in self.rows),
                                 ptr(desc))
        check(rc, hd, "hps_pack_blocks")
        self.rows = []


def _addr(t, offset_elems=0):
    return t.data_ptr() + offset_elems * t.element_size()


# ---------------------------------------------------------------------------
# stages
# ---------------------------------------------------------------------------
def local_solve_stage_adaptive_2D_DtN(pde_problem, device=None, host_device=None):
    """Leaf DtN maps for leaves of different sizes (``local_solve/_adaptive_2D_DtN.py:13-117``).
    Returns Y (n_leaves, p^2, 4q), T (n_leaves, 4q, 4q), v (n_leaves, p^2), h (n_leaves, 4q)
    (a trailing source axis is kept for multi-source problems)."""
    from ._stages import leaf_solve
    from .local_solve import _compute_dtype, _resolve, _source_3d

    dev = _resolve(device)
    dom = pde_problem.domain
    dt = _compute_dtype(pde_problem)
    src, multi = _source_3d(pde_problem, dt, dev)
    half = np.array([(lf.xmax - lf.xmin) / 2 for lf in get_all_leaves(dom.root)], dtype=np.float64)
    res = leaf_solve(pde_problem, HPS_KIND_DTN, dt, dev, source=src, leaf_half_sides=half)
    v, h = res.v, res.h
    if not multi:
        v, h = v[..., 0], h[..., 0]
    return res.Y, res.T, v, h


_STREAMS: dict = {}


def _side_streams(dev, n=4):
    """A small pool of default-priority streams on which the independent merge
    groups of one tree level run concurrently (their kernels are small and
    latency-bound)."""
    key = str(dev)
    if key not in _STREAMS:
        _STREAMS[key] = [torch.cuda.Stream(device=dev) for _ in range(n)]
    return _STREAMS[key]


def _run_level(dev, items, body):
    """body(*item) for every item, spread over the side streams; the current
    stream waits for all of them before returning, and they wait for it
    before starting."""
    main = torch.cuda.current_stream(dev)
    if len(items) <= 1:
        for it in items:
            body(*it)
        return
    ss = _side_streams(dev)
    for st in ss:
        st.wait_stream(main)
    for i, it in enumerate(items):
        with torch.cuda.stream(ss[i % len(ss)]):
            body(*it)
    for st in ss:
        main.wait_stream(st)


def _merge_group(plan, plist, out_list):
    """Pack, compress and merge the children of the parents in plist (one
    signature); appends (plan, plist, merge outputs)."""
    from ._stages import merge_level

    nc = plan.nc
    T0 = plist[0].children[0].data.T
    dev, dt = T0.device, T0.dtype
    nsrc = plist[0].children[0].data.h.shape[-1]
    nm, nmax = len(plist), plan.nmax
    Tin = torch.empty((nm * nc, nmax, nmax), dtype=dt, device=dev)
    Hin = torch.empty((nm * nc, nmax, nsrc), dtype=dt, device=dev)
    # compressed positions are staged per position, then compressed by two
    # batched products (R T) P and one R h
    stage = {ch: (torch.empty((nm, plan.n[ch], plan.n[ch]), dtype=dt, device=dev),
                  torch.empty((nm, plan.n[ch], nsrc), dtype=dt, device=dev))
             for ch in range(nc) if plan.R[ch] is not None}
    pk = _Packer(dt, dev)
    for k, nd in enumerate(plist):
        for ch, child in enumerate(nd.children):
            n_c = plan.n[ch]
            Tc, hc = child.data.T, child.data.h
            if ch in stage:
                pk.add(Tc, _addr(stage[ch][0], k * n_c * n_c), n_c, n_c, n_c, n_c)
                pk.add(hc, _addr(stage[ch][1], k * n_c * nsrc), n_c, nsrc, nsrc, nsrc)
            else:
                slot = nc * k + ch
                pk.add(Tc, _addr(Tin, slot * nmax * nmax), n_c, n_c, n_c, nmax)
                pk.add(hc, _addr(Hin, slot * nmax * nsrc), n_c, nsrc, nsrc, nsrc)
    pk.run()
    for ch, (Ts, Hs) in stage.items():
        n_c, m_c = plan.n[ch], plan.m[ch]
        R, P = plan.ops(ch, dev, dt)
        tmp = torch.empty((nm, m_c, n_c), dtype=dt, device=dev)
        _gemm_strided(m_c, n_c, n_c, ptr(R), n_c, 0, ptr(Ts), n_c, n_c * n_c, ptr(tmp), n_c, m_c * n_c, nm, dt)
        _gemm_strided(m_c, m_c, n_c, ptr(tmp), n_c, m_c * n_c, ptr(P), m_c, 0,
                      ctypes.c_void_p(_addr(Tin, ch * nmax * nmax)), nmax, nc * nmax * nmax, nm, dt)
        _gemm_strided(m_c, nsrc, n_c, ptr(R), n_c, 0, ptr(Hs), nsrc, n_c * nsrc,
                      ctypes.c_void_p(_addr(Hin, ch * nmax * nsrc)), nsrc, nc * nmax * nsrc, nm, dt)
    out = merge_level(Tin, Hin, plan)
    for k, nd in enumerate(plist):
        nd.data.T, nd.data.h = out.T[k], out.h[k]
        nd.data.S, nd.data.g_tilde = out.S[k], out.gt[k]
    out_list.append((plan, plist, out))


def adaptive_merge_stage(pde_problem, T_leaf, h_leaf, signature, make_plan) -> None:
    """Merges from the deepest level to the root for 2D or 3D adaptive trees.
    Groups of one level (one per signature) run concurrently; the pivots are
    checked once at the end."""
    dom = pde_problem.domain
    root, q = dom.root, dom.q
    if T_leaf is not None:
        for i, lf in enumerate(get_all_leaves(root)):
            lf.data.T = T_leaf[i]
            lf.data.h = h_leaf[i] if h_leaf[i].dim() == 2 else h_leaf[i][:, None]
    cache = pde_problem.__dict__.setdefault("_adaptive_plans", {})
    dev = get_all_leaves(root)[0].data.T.device
    dt = get_all_leaves(root)[0].data.T.dtype
    levels = []
    for level in range(get_depth(root), 0, -1):
        parents = [nd for nd in get_nodes_at_level(root, level - 1) if len(nd.children)]
        groups = OrderedDict()
        for nd in parents:
            groups.setdefault(signature(nd, q), []).append(nd)
        items = []
        for sig, plist in groups.items():
            plan = cache.get(sig)
            if plan is None:
                plan = cache[sig] = make_plan(sig)
            # device tables and operators on the current stream, before the
            # side streams read them
            plan.device(dev)
            for ch in range(plan.nc):
                if plan.R[ch] is not None:
                    plan.ops(ch, dev, dt)
            items.append((plan, plist, levels))
        _run_level(dev, items, _merge_group)
    if levels:
        bad = torch.cat([out.info for _, _, out in levels])
        _lib.raise_on_info(bad, "adaptive merge")
    pde_problem._adaptive_levels = levels


def merge_stage_adaptive_2D_DtN(pde_problem, T_leaf=None, h_leaf=None) -> None:
    """Merges from the deepest level to the root (``merge/_adaptive_2D_DtN.py:25-110``);
    each merged node keeps data.T, data.h, data.S, data.g_tilde (device tensors,
    h and g_tilde with a source axis)."""
    L21, L12 = pde_problem.L_2f1, pde_problem.L_1f2
    q = pde_problem.domain.q
    adaptive_merge_stage(pde_problem, T_leaf, h_leaf, _signature, lambda sig: _Plan(sig[0], sig[1], q, L21, L12))


def down_pass_adaptive_2D_DtN(pde_problem, boundary_data):
    """Root Dirichlet data (a list of the four sides' -- six faces' in 3D --
    samples, or one array)
    down to the leaves (``down_pass/_adaptive_2D_DtN.py:16-98``).  Returns u
    (n_leaves, p^2) (with a trailing source axis for multi-source data)."""
    from ._stages import down_level, leaf_apply

    dom = pde_problem.domain
    root = dom.root
    Y, v = pde_problem.Y, pde_problem.v
    dev, dt = Y.device, Y.dtype
    if isinstance(boundary_data, (list, tuple)):
        g = torch.cat([torch.as_tensor(np.asarray(b) if not isinstance(b, torch.Tensor) else b).to(dev, dt)
                       .reshape(len(b), -1) for b in boundary_data])
        single = all(np.ndim(b) == 1 for b in boundary_data)
    else:
        g = torch.as_tensor(boundary_data).to(dev, dt)
        single = g.dim() == 1
        g = g.reshape(g.shape[0], -1)
    if g.shape[0] != dom.boundary_points.shape[0]:
        raise ValueError(f"boundary data has {g.shape[0]} points, the root boundary {dom.boundary_points.shape[0]}")
    root.data.g = g.contiguous()
    nsrc = g.shape[1]
    by_level = OrderedDict()
    for plan, plist, out in reversed(pde_problem._adaptive_levels):
        by_level.setdefault(plist[0].depth, []).append((plan, plist, out))

    def group_down(plan, plist, out):
        nm, nmax, ne = len(plist), plan.nmax, plan.ne
        gt = out.gt
        if gt.shape[-1] != nsrc:
            if gt.shape[-1] != 1:
                raise ValueError("boundary data and the build's sources disagree in count")
            gt = gt.expand(-1, -1, nsrc).contiguous()
        gin = torch.empty((nm, ne, nsrc), dtype=dt, device=dev)
        pk = _Packer(dt, dev)
        for k, nd in enumerate(plist):
            pk.add(nd.data.g, _addr(gin, k * ne * nsrc), ne, nsrc, nsrc, nsrc)
        pk.run()
        gout = down_level(out.S, gt, gin, plan)
        nc = plan.nc
        for ch in range(nc):
            n_c, m_c = plan.n[ch], plan.m[ch]
            if plan.P[ch] is None:
                for k, nd in enumerate(plist):
                    nd.children[ch].data.g = gout[nc * k + ch, :n_c]
            else:
                _, P = plan.ops(ch, dev, dt)
                gc = torch.empty((nm, n_c, nsrc), dtype=dt, device=dev)
                _gemm_strided(n_c, nsrc, m_c, ptr(P), m_c, 0, ctypes.c_void_p(_addr(gout, ch * nmax * nsrc)), nsrc,
                              nc * nmax * nsrc, ptr(gc), nsrc, n_c * nsrc, nm, dt)
                for k, nd in enumerate(plist):
                    nd.children[ch].data.g = gc[k]

    for items in by_level.values():
        _run_level(dev, items, group_down)
    leaves = get_all_leaves(root)
    nb = leaves[0].data.g.shape[0]
    G = torch.empty((len(leaves), nb, nsrc), dtype=dt, device=dev)
    pk = _Packer(dt, dev)
    for i, lf in enumerate(leaves):
        pk.add(lf.data.g, _addr(G, i * nb * nsrc), nb, nsrc, nsrc, nsrc)
    pk.run()
    vv = v if v.dim() == 3 else v[..., None]
    if vv.shape[-1] != nsrc:
        vv = vv.expand(-1, -1, nsrc)
    u = leaf_apply(Y.contiguous(), G, vv.contiguous())
    return u[..., 0] if single and nsrc == 1 else u
