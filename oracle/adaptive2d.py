"""ORACLE (test infrastructure only) -- adaptive 2D DtN restatement.

Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` may import this package; it is the checker, never the product.

Restates, for quadtrees whose neighbouring leaves differ by at most one level:

* the adaptive refinement with level restriction
  (``_adaptive_discretization_2D.py:35-327``) and its panel norms
  (``:362-407``);
* the per-leaf DtN solve with operators scaled by each leaf's side
  (``local_solve/_adaptive_2D_DtN.py:13-172``);
* 4:1 merges of children with different boundary discretisations: panel
  pairs of the finer side of an interface are compressed with the
  two-panels-to-one interpolation L_1f2 (rows) and the one-to-two
  interpolation L_2f1 (columns) before the Schur complement
  (``merge/_adaptive_2D_DtN.py:25-433``, ``merge/_utils_adaptive_2D_DtN.py``);
* the down pass, which refines compressed interface data with L_2f1
  (``down_pass/_adaptive_2D_DtN.py:16-263``).

Works on any node objects with ``xmin, xmax, ymin, ymax, depth, children``.
Pinned by (tests): identity with the uniform restatement (``oracle.hps2d``)
on uniform trees, exact reproduction of low-degree polynomial solutions on
non-uniform trees, the panel-norm closed forms of the reference's tests, and
the level-restriction invariant of the refined trees.
"""

from __future__ import annotations

import numpy as np

from . import hps2d
from . import spectral as osp


# ---------------------------------------------------------------------------
# trees
# ---------------------------------------------------------------------------
class Node:
    def __init__(self, xmin, xmax, ymin, ymax, depth=0):
        self.xmin, self.xmax, self.ymin, self.ymax = xmin, xmax, ymin, ymax
        self.depth = depth
        self.children = []


def split(node):
    """Children SW, SE, NE, NW (``_discretization_tree_operations_2D.py:8-47``)."""
    xm, ym = (node.xmin + node.xmax) / 2, (node.ymin + node.ymax) / 2
    d = node.depth + 1
    node.children = [
        Node(node.xmin, xm, node.ymin, ym, d),
        Node(xm, node.xmax, node.ymin, ym, d),
        Node(xm, node.xmax, ym, node.ymax, d),
        Node(node.xmin, xm, ym, node.ymax, d),
    ]


def leaves(node):
    if not len(node.children):
        return [node]
    return [lf for c in node.children for lf in leaves(c)]


def bounds(node):
    return (node.xmin, node.xmax, node.ymin, node.ymax)


def side_counts(node, q):
    """Gauss points on sides S, E, N, W = q x the leaves touching each side."""
    lv = leaves(node)
    return (
        q * sum(lf.ymin == node.ymin for lf in lv),
        q * sum(lf.xmax == node.xmax for lf in lv),
        q * sum(lf.ymax == node.ymax for lf in lv),
        q * sum(lf.xmin == node.xmin for lf in lv),
    )


def side_leaves(node):
    """Leaves along sides S (W->E), E (S->N), N (E->W), W (N->S)."""
    lv = leaves(node)
    s = sorted([lf for lf in lv if lf.ymin == node.ymin], key=lambda lf: lf.xmin)
    e = sorted([lf for lf in lv if lf.xmax == node.xmax], key=lambda lf: lf.ymin)
    n = sorted([lf for lf in lv if lf.ymax == node.ymax], key=lambda lf: -lf.xmin)
    w = sorted([lf for lf in lv if lf.xmin == node.xmin], key=lambda lf: -lf.ymin)
    return s, e, n, w


def boundary_points(root, q):
    """Root boundary Gauss points in the order above (each panel's points
    running with the walk)."""
    g = osp.gauss_nodes(q)
    s, e, n, w = side_leaves(root)
    out = []
    for lf in s:
        out += [(0.5 * (lf.xmin + lf.xmax) + 0.5 * (lf.xmax - lf.xmin) * t, root.ymin) for t in g]
    for lf in e:
        out += [(root.xmax, 0.5 * (lf.ymin + lf.ymax) + 0.5 * (lf.ymax - lf.ymin) * t) for t in g]
    for lf in n:
        out += [(0.5 * (lf.xmin + lf.xmax) - 0.5 * (lf.xmax - lf.xmin) * t, root.ymax) for t in g]
    for lf in w:
        out += [(root.xmin, 0.5 * (lf.ymin + lf.ymax) - 0.5 * (lf.ymax - lf.ymin) * t) for t in g]
    return np.array(out)


def interior_points(root, p):
    pts = []
    _, X, Y = osp.leaf_order_2D(p)
    for lf in leaves(root):
        cx, cy = (lf.xmin + lf.xmax) / 2, (lf.ymin + lf.ymax) / 2
        hx, hy = (lf.xmax - lf.xmin) / 2, (lf.ymax - lf.ymin) / 2
        pts.append(np.stack([cx + hx * X, cy + hy * Y], -1))
    return np.array(pts)


def level_restricted(root):
    """Every pair of leaves sharing an edge segment differs by <= 1 level."""
    lv = leaves(root)
    for a in lv:
        for b in lv:
            if abs(a.depth - b.depth) <= 1:
                continue
            touch_x = (a.xmax == b.xmin or b.xmax == a.xmin) and min(a.ymax, b.ymax) > max(a.ymin, b.ymin)
            touch_y = (a.ymax == b.ymin or b.ymax == a.ymin) and min(a.xmax, b.xmax) > max(a.xmin, b.xmin)
            if touch_x or touch_y:
                return False
    return True


# ---------------------------------------------------------------------------
# adaptive refinement
# ---------------------------------------------------------------------------
def _cc_weights(n, a, b):
    """Clenshaw-Curtis weights on the ascending Chebyshev points of [a, b]."""
    x = osp.cheb_points(n)
    # integrate the Lagrange basis exactly with Gauss-Legendre of enough points
    g, w = np.polynomial.legendre.leggauss(n + 2)
    L = osp.lagrange_matrix(x, g)
    return (w @ L) * (b - a) / 2


def panel_sq_l2(f_vals, b, p):
    """Squared L2 norm of f on the panel b from its leaf-ordered samples
    (``_adaptive_discretization_2D.py:384-407``)."""
    perm, _, _ = osp.leaf_order_2D(p)
    wx = _cc_weights(p, b[0], b[1])
    wy = _cc_weights(p, b[2], b[3])
    # leaf point k sits at grid (i, j) of meshgrid(x, flip(y), "ij")
    W = np.outer(wx, wy[::-1]).reshape(-1)[perm]
    return float(np.sum(f_vals**2 * W))


def _refine_operator(p):
    """Interpolation from a leaf's p^2 grid to its four children's grids
    (children SW, SE, NE, NW, each in leaf order) on [-1, 1]^2."""
    _, X, Y = osp.leaf_order_2D(p)
    c = osp.cheb_points(p)
    rows = []
    for (cx, cy) in ((-0.5, -0.5), (0.5, -0.5), (0.5, 0.5), (-0.5, 0.5)):
        xs, ys = cx + 0.5 * X, cy + 0.5 * Y
        Lx = osp.lagrange_matrix(c, xs)
        Ly = osp.lagrange_matrix(c, ys)
        # parent point k is at (X[k], Y[k]): weight Lx[:, ix(k)] * Ly[:, iy(k)]
        ix = np.array([int(np.argmin(np.abs(c - v))) for v in X])
        iy = np.array([int(np.argmin(np.abs(c - v))) for v in Y])
        rows.append(Lx[:, ix] * Ly[:, iy])
    return np.concatenate(rows, axis=0)


def _panel_pts(b, p):
    _, X, Y = osp.leaf_order_2D(p)
    cx, cy = (b[0] + b[1]) / 2, (b[2] + b[3]) / 2
    hx, hy = (b[1] - b[0]) / 2, (b[3] - b[2]) / 2
    return np.stack([cx + hx * X, cy + hy * Y], -1)


def _children_pts(b, p):
    xm, ym = (b[0] + b[1]) / 2, (b[2] + b[3]) / 2
    kids = [(b[0], xm, b[2], ym), (xm, b[1], b[2], ym), (xm, b[1], ym, b[3]), (b[0], xm, ym, b[3])]
    return np.concatenate([_panel_pts(k, p) for k in kids]), kids


def _neighbour_boxes(node, root):
    """Same-size boxes W, E, S, N of node inside the root
    (``_adaptive_discretization_2D.py:212-274``)."""
    s = node.xmax - node.xmin
    out = []
    if node.xmin - s >= root.xmin:
        out.append((node.xmin - s, node.xmin, node.ymin, node.ymax))
    if node.xmax + s <= root.xmax:
        out.append((node.xmax, node.xmax + s, node.ymin, node.ymax))
    if node.ymin - s >= root.ymin:
        out.append((node.xmin, node.xmax, node.ymin - s, node.ymin))
    if node.ymax + s <= root.ymax:
        out.append((node.xmin, node.xmax, node.ymax, node.ymax + s))
    return out


def _ensure_box(node, box):
    """Make the box (one quadtree cell) exist by splitting its ancestor cell
    if that is a leaf; return the node split, or None
    (``_adaptive_discretization_2D.py:277-359``)."""
    want = box[1] - box[0]
    side = node.xmax - node.xmin
    if np.isclose(2 * want, side, rtol=1e-5, atol=1e-8):
        if not node.children:
            split(node)
            return node
        return None
    xm, ym = (node.xmin + node.xmax) / 2, (node.ymin + node.ymax) / 2
    k = (0 if box[2] < ym else 3) if box[0] < xm else (1 if box[2] < ym else 2)
    child = node.children[k]
    if np.isclose(side, 4 * want, rtol=1e-5, atol=1e-8):
        if not child.children:
            split(child)
            return child
        return None
    if side < 2 * want:
        raise ValueError("requested box is larger than the node")
    return _ensure_box(child, box)


def refine(root, f, tol, p, restrict=True, l2=False):
    """generate_adaptive_mesh_level_restriction_2D on an oracle tree."""
    R = _refine_operator(p)
    if l2:
        if not root.children:
            split(root)
        gnorm = sum(panel_sq_l2(f(_panel_pts(bounds(lf), p)), bounds(lf), p) for lf in leaves(root))
        tol = tol**2
    else:
        pts, _ = _children_pts(bounds(root), p)
        gnorm = float(np.max(f(pts)))
    queue = leaves(root)
    while queue:
        ok, norms = [], []
        for nd in queue:
            b = bounds(nd)
            f0 = f(_panel_pts(b, p))
            pts1, kids = _children_pts(b, p)
            f1 = f(pts1)
            diff = R @ f0 - f1
            if l2:
                m = p * p
                err = sum(panel_sq_l2(diff[i * m:(i + 1) * m], kids[i], p) for i in range(4))
                ok.append(err / gnorm < tol)
                norms.append(0.0)
            else:
                c = max(gnorm, float(np.max(np.abs(f1))))
                ok.append(float(np.max(np.abs(diff))) / c < tol)
                norms.append(c)
        if not l2:
            gnorm = max(norms)
        nxt = []
        for nd, good in zip(queue, ok):
            if good:
                continue
            if not nd.children:
                split(nd)
            nxt.extend(nd.children)
            if restrict:
                check = [nd]
                while check:
                    cur = check.pop()
                    for box in _neighbour_boxes(cur, root):
                        made = _ensure_box(root, box)
                        if made is not None:
                            check.append(made)
                            nxt.extend(made.children)
        queue = nxt


# ---------------------------------------------------------------------------
# leaf stage
# ---------------------------------------------------------------------------
def leaf_stage(root, p, q, coeffs, source):
    """Per-leaf DtN maps; coeffs: dict name -> (n_leaves, p^2); source
    (n_leaves, p^2[, nsrc]).  Returns Y, T, v, h with a source axis."""
    lv = leaves(root)
    f = source if source.ndim == 3 else source[..., None]
    Ys, Ts, vs, hs = [], [], [], []
    ops_of = {}
    for i, lf in enumerate(lv):
        half = (lf.xmax - lf.xmin) / 2
        if half not in ops_of:
            ops_of[half] = osp.leaf_ops_2D(p, q, half)
        ops = ops_of[half]
        A = hps2d.assemble_diff_operator({k: v[i:i + 1] for k, v in coeffs.items()}, ops)
        Y, T, v, h = hps2d.leaf_dtn(A, f[i:i + 1], ops["P"], ops["Q"])
        Ys.append(Y[0])
        Ts.append(T[0])
        vs.append(v[0])
        hs.append(h[0])
    return np.array(Ys), np.array(Ts), np.array(vs), np.array(hs)


# ---------------------------------------------------------------------------
# merges
# ---------------------------------------------------------------------------
def projection_ops(q):
    """L_2f1 (2q x q) and L_1f2 (q x 2q) (``_precompute_operators_2D.py:269-300``)."""
    g = osp.gauss_nodes(q)
    g2 = np.concatenate([0.5 * g - 0.5, 0.5 * g + 0.5])
    return osp.lagrange_matrix(g, g2), osp.lagrange_matrix(g2, g)


def _walk(first, second, inner_end, outward_sign):
    """Two-pointer walk from the outer boundary inwards over two lists of
    leaves along an interface; inner_end(leaf) is the leaf's coordinate
    nearest the merged node's centre.  Returns the panel compression flags."""
    out1, out2 = [], []
    i = j = 0
    while i < len(first) and j < len(second):
        e1, e2 = inner_end(first[i]), inner_end(second[j])
        if e1 == e2:
            out1.append(False)
            out2.append(False)
            i += 1
            j += 1
        elif (e1 - e2) * outward_sign > 0:
            # first's leaf ends nearer the outside: its next two panels are finer
            out1 += [True, True]
            out2.append(False)
            i += 2
            j += 1
        else:
            out2 += [True, True]
            out1.append(False)
            j += 2
            i += 1
    return out1, out2


def compression_lists(a, b, c, d):
    """Flags for a5, b5, b6, c6, c7, d7, d8, a8, each outside -> inside
    (``merge/_utils_adaptive_2D_DtN.py:10-250``)."""
    la, lb, lc, ld = leaves(a), leaves(b), leaves(c), leaves(d)
    a5 = sorted([x for x in la if x.xmax == a.xmax], key=lambda x: x.ymin)
    b5 = sorted([x for x in lb if x.xmin == b.xmin], key=lambda x: x.ymin)
    b6 = sorted([x for x in lb if x.ymax == b.ymax], key=lambda x: -x.xmin)
    c6 = sorted([x for x in lc if x.ymin == c.ymin], key=lambda x: -x.xmin)
    c7 = sorted([x for x in lc if x.xmin == c.xmin], key=lambda x: -x.ymin)
    d7 = sorted([x for x in ld if x.xmax == d.xmax], key=lambda x: -x.ymin)
    d8 = sorted([x for x in ld if x.ymin == d.ymin], key=lambda x: x.xmin)
    a8 = sorted([x for x in la if x.ymax == a.ymax], key=lambda x: x.xmin)
    f_a5, f_b5 = _walk(a5, b5, lambda x: x.ymax, -1)  # inwards = +y
    f_b6, f_c6 = _walk(b6, c6, lambda x: x.xmin, +1)  # inwards = -x
    f_c7, f_d7 = _walk(c7, d7, lambda x: x.ymin, +1)  # inwards = -y
    f_d8, f_a8 = _walk(d8, a8, lambda x: x.xmax, -1)  # inwards = +x
    return f_a5, f_b5, f_b6, f_c6, f_c7, f_d7, f_d8, f_a8


def _compress_rows(M, L12, flags, q):
    out, i = [], 0
    while i < len(flags):
        if flags[i]:
            out.append(L12 @ M[q * i:q * (i + 2)])
            i += 2
        else:
            out.append(M[q * i:q * (i + 1)])
            i += 1
    return np.concatenate(out, axis=0) if out else M[:0]


def _compress_cols(M, L21, flags, q):
    return _compress_rows(M.T, L21.T, flags, q).T


def _child_blocks(T, h, idx, flags, L21, L12, q):
    """Blocks T[idx_i][:, idx_j] and h[idx_i] with the panels flagged in
    flags[i] compressed (rows with L_1f2, columns with L_2f1)."""
    hs = [_compress_rows(h[ix], L12, fl, q) for ix, fl in zip(idx, flags)]
    Ts = [[_compress_cols(_compress_rows(T[np.ix_(ii, jj)], L12, fi, q), L21, fj, q) for jj, fj in zip(idx, flags)]
          for ii, fi in zip(idx, flags)]
    return hs, Ts


def _index_sets(n_sides, which):
    n0, n1, n2, n3 = n_sides
    r = np.arange(n0 + n1 + n2 + n3)
    if which == "a":  # ext 1, interface 5, interface 8
        return np.concatenate([r[n0 + n1 + n2:], r[:n0]]), r[n0:n0 + n1], r[n0 + n1:n0 + n1 + n2][::-1]
    if which == "b":  # ext 2, interface 6, interface 5
        return r[:n0 + n1], r[n0 + n1:n0 + n1 + n2], r[n0 + n1 + n2:][::-1]
    if which == "c":  # interface 6, ext 3, interface 7
        return r[:n0][::-1], r[n0:n0 + n1 + n2], r[n0 + n1 + n2:]
    # d: interface 8, interface 7, ext 4
    return r[:n0], r[n0:n0 + n1][::-1], r[n0 + n1:]


def quad_merge(node, Ts, hs, q):
    """Merge of node's four children (maps Ts, data hs (n, nsrc)):
    S, T, h_out, g_tilde."""
    a, b, c, d = node.children
    L21, L12 = projection_ops(q)
    fl = compression_lists(a, b, c, d)
    nsides = [side_counts(x, q) for x in (a, b, c, d)]
    ext_a = [False] * (len(_index_sets(nsides[0], "a")[0]) // q)
    ext_b = [False] * (len(_index_sets(nsides[1], "b")[0]) // q)
    ext_c = [False] * (len(_index_sets(nsides[2], "c")[1]) // q)
    ext_d = [False] * (len(_index_sets(nsides[3], "d")[2]) // q)
    (ha1, ha5, ha8), ((a11, a15, a18), (a51, a55, a58), (a81, a85, a88)) = _child_blocks(
        Ts[0], hs[0], _index_sets(nsides[0], "a"), (ext_a, fl[0], fl[7]), L21, L12, q)
    (hb2, hb6, hb5), ((b22, b26, b25), (b62, b66, b65), (b52, b56, b55)) = _child_blocks(
        Ts[1], hs[1], _index_sets(nsides[1], "b"), (ext_b, fl[2], fl[1]), L21, L12, q)
    (hc6, hc3, hc7), ((c66, c63, c67), (c36, c33, c37), (c76, c73, c77)) = _child_blocks(
        Ts[2], hs[2], _index_sets(nsides[2], "c"), (fl[3], ext_c, fl[4]), L21, L12, q)
    (hd8, hd7, hd4), ((d88, d87, d84), (d78, d77, d74), (d48, d47, d44)) = _child_blocks(
        Ts[3], hs[3], _index_sets(nsides[3], "d"), (fl[6], fl[5], ext_d), L21, L12, q)
    n1, n2, n3, n4 = a11.shape[0], b22.shape[0], c33.shape[0], d44.shape[0]
    n5, n6, n7, n8 = a55.shape[0], b66.shape[0], c77.shape[0], d88.shape[0]
    Z = lambda r, k: np.zeros((r, k))
    B = np.block([[a15, Z(n1, n6), Z(n1, n7), a18], [b25, b26, Z(n2, n7), Z(n2, n8)],
                  [Z(n3, n5), c36, c37, Z(n3, n8)], [Z(n4, n5), Z(n4, n6), d47, d48]])
    C = np.block([[a51, b52, Z(n5, n3), Z(n5, n4)], [Z(n6, n1), b62, c63, Z(n6, n4)],
                  [Z(n7, n1), Z(n7, n2), c73, d74], [a81, Z(n8, n2), Z(n8, n3), d84]])
    D = np.block([[a55 + b55, b56, Z(n5, n7), a58], [b65, b66 + c66, c67, Z(n6, n8)],
                  [Z(n7, n5), c76, c77 + d77, d78], [a85, Z(n8, n6), d87, d88 + a88]])
    h_int = np.concatenate([ha5 + hb5, hb6 + hc6, hc7 + hd7, hd8 + ha8])
    h_ext = np.concatenate([ha1, hb2, hc3, hd4])
    Dinv = np.linalg.inv(D)
    S = -Dinv @ C
    T = B @ S
    o = 0
    for Ablk in (a11, b22, c33, d44):
        k = Ablk.shape[0]
        T[o:o + k, o:o + k] += Ablk
        o += k
    g_tilde = -Dinv @ h_int
    h_out = h_ext + B @ g_tilde
    m = nsides[0][3]
    return np.roll(S, -m, 1), np.roll(np.roll(T, -m, 0), -m, 1), np.roll(h_out, -m, 0), g_tilde


def merge_stage(root, T_leaf, h_leaf, q):
    """Merges bottom-up; returns dicts id(node) -> (S, g_tilde) and the root
    (T, h).  T_leaf/h_leaf in leaf order."""
    Tm, hm = {}, {}
    for lf, T, h in zip(leaves(root), T_leaf, h_leaf):
        Tm[id(lf)], hm[id(lf)] = T, h
    Sg = {}

    def rec(nd):
        if not nd.children:
            return
        for ch in nd.children:
            rec(ch)
        S, T, h, g = quad_merge(nd, [Tm[id(c)] for c in nd.children], [hm[id(c)] for c in nd.children], q)
        Tm[id(nd)], hm[id(nd)], Sg[id(nd)] = T, h, (S, g)

    rec(root)
    return Sg, Tm[id(root)], hm[id(root)]


# ---------------------------------------------------------------------------
# down pass
# ---------------------------------------------------------------------------
def _decompress(g_int, f1, f2, L21, q, start):
    """Split compressed interface data into the two children's versions."""
    o1, o2 = [], []
    i = j = 0
    k = start
    while i < len(f1) and j < len(f2):
        panel = g_int[k:k + q]
        if f1[i]:
            o1.append(L21 @ panel)
            i += 2
        else:
            o1.append(panel)
            i += 1
        if f2[j]:
            o2.append(L21 @ panel)
            j += 2
        else:
            o2.append(panel)
            j += 1
        k += q
    return np.concatenate(o1), np.concatenate(o2), k


def down_pass(root, bdry, Sg, Y, v, q):
    """bdry: root boundary data (n_bdry, nsrc).  Returns u (n_leaves, p^2, nsrc)."""
    L21, _ = projection_ops(q)
    g_leaf = {}

    def rec(nd, g):
        if not nd.children:
            g_leaf[id(nd)] = g
            return
        a, b, c, d = nd.children
        S, gt = Sg[id(nd)]
        gi = S @ g + gt
        fl = compression_lists(a, b, c, d)
        na, nb, nc, nd_ = (side_counts(x, q) for x in (a, b, c, d))
        n0, n1, n2 = na[0] + nb[0], nb[1] + nc[1], nc[2] + nd_[2]
        s0, s1, s2, s3 = g[:n0], g[n0:n0 + n1], g[n0 + n1:n0 + n1 + n2], g[n0 + n1 + n2:]
        ga5, gb5, k = _decompress(gi, fl[0], fl[1], L21, q, 0)
        gb6, gc6, k = _decompress(gi, fl[2], fl[3], L21, q, k)
        gc7, gd7, k = _decompress(gi, fl[4], fl[5], L21, q, k)
        gd8, ga8, k = _decompress(gi, fl[6], fl[7], L21, q, k)
        rec(a, np.concatenate([s0[:na[0]], ga5, ga8[::-1], s3[nd_[3]:]]))
        rec(b, np.concatenate([s0[na[0]:], s1[:nb[1]], gb6, gb5[::-1]]))
        rec(c, np.concatenate([gc6[::-1], s1[nb[1]:], s2[:nc[2]], gc7]))
        rec(d, np.concatenate([gd8, gd7[::-1], s2[nc[2]:], s3[:nd_[3]]]))

    rec(root, bdry)
    lv = leaves(root)
    return np.array([Y[i] @ g_leaf[id(lf)] + v[i] for i, lf in enumerate(lv)])


def solve(root, p, q, coeffs, source, bdry):
    """Whole adaptive DtN build + solve: u (n_leaves, p^2, nsrc) and root T."""
    Y, T, v, h = leaf_stage(root, p, q, coeffs, source)
    Sg, Troot, _ = merge_stage(root, T, h, q)
    b = bdry if bdry.ndim == 2 else bdry[:, None]
    return down_pass(root, b, Sg, Y, v, q), Troot
