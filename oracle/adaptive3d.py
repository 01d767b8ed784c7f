"""ORACLE (test infrastructure only) -- adaptive 3D DtN restatement.

Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` may import this package; it is the checker, never the product.

Restates, for octrees whose face-neighbouring leaves differ by at most one
level: the per-leaf DtN solve with operators scaled by each leaf's side
(``local_solve/_adaptive_3D_DtN.py``); 8:1 merges in which the finer side of
an interface has each group of four face panels compressed to one with the
coarsening operator L_1f4 (rows) and the refining operator L_4f1 (columns)
(``merge/_adaptive_3D_DtN.py``, ``merge/_utils_adaptive_3D_DtN.py``); the
down pass, which refines compressed interface data with L_4f1
(``down_pass/_adaptive_3D_DtN.py``); and the refinement with level
restriction (``_adaptive_discretization_3D.py``).

Faces 0..5 = x-, x+, y-, y+, z-, z+; a face's panels run in the recursive quad
order of the merged faces (``EXT_FACES``), each panel's q^2 Gauss points with
the first face coordinate slowest -- the uniform conventions of
``oracle.hps3d``.  Pinned by (tests): identity with ``oracle.hps3d`` on
uniform trees, exact low-degree polynomial solutions on non-uniform trees,
the projection operators' exactness, and the level-restriction invariant.
"""

from __future__ import annotations

import numpy as np

from . import hps2d, hps3d
from . import spectral as osp

IFACES = hps3d.IFACES
EXT_FACES = hps3d.EXT_FACES


class Node:
    def __init__(self, b, depth=0):
        self.xmin, self.xmax, self.ymin, self.ymax, self.zmin, self.zmax = b
        self.depth = depth
        self.children = []


def bounds(nd):
    return (nd.xmin, nd.xmax, nd.ymin, nd.ymax, nd.zmin, nd.zmax)


def split(nd):
    nd.children = [Node(b, nd.depth + 1) for b in hps3d.oct_children(bounds(nd))]


def leaves(nd):
    if not nd.children:
        return [nd]
    return [lf for c in nd.children for lf in leaves(c)]


def face_leaves(nd, f):
    """Leaves on face f of nd, in the face's panel order."""
    if not nd.children:
        return [nd]
    return [lf for c in EXT_FACES[f] for lf in face_leaves(nd.children[c], f)]


def face_counts(nd, q):
    return tuple(q * q * len(face_leaves(nd, f)) for f in range(6))


def side(nd):
    return nd.xmax - nd.xmin


def interior_points(root, p):
    _, pts = hps3d.leaf_order_3D(p)
    out = []
    for lf in leaves(root):
        b = bounds(lf)
        c = np.array([(b[0] + b[1]) / 2, (b[2] + b[3]) / 2, (b[4] + b[5]) / 2])
        h = np.array([(b[1] - b[0]) / 2, (b[3] - b[2]) / 2, (b[5] - b[4]) / 2])
        out.append(c + h * pts)
    return np.array(out)


def boundary_points(root, q):
    b = bounds(root)
    lims = [(b[0], b[1]), (b[2], b[3]), (b[4], b[5])]
    out = []
    for f, (ax, sg, a1, a2) in enumerate(hps3d.FACES):
        for lf in face_leaves(root, f):
            lb = bounds(lf)
            ll = [(lb[0], lb[1]), (lb[2], lb[3]), (lb[4], lb[5])]
            fp = hps3d._face_patch_points(*ll[a1], *ll[a2], q)
            xyz = np.zeros((q * q, 3))
            xyz[:, ax] = lims[ax][1] if sg > 0 else lims[ax][0]
            xyz[:, a1] = fp[:, 0]
            xyz[:, a2] = fp[:, 1]
            out.append(xyz)
    return np.concatenate(out)


def level_restricted(root):
    lv = leaves(root)
    for a in lv:
        for b in lv:
            if abs(a.depth - b.depth) <= 1:
                continue
            A, B = bounds(a), bounds(b)
            for ax in range(3):
                if A[2 * ax + 1] == B[2 * ax] or B[2 * ax + 1] == A[2 * ax]:
                    others = [k for k in range(3) if k != ax]
                    if all(min(A[2 * k + 1], B[2 * k + 1]) > max(A[2 * k], B[2 * k]) for k in others):
                        return False
    return True


# ---------------------------------------------------------------------------
# projection operators (_precompute_operators_3D.py:323-409)
# ---------------------------------------------------------------------------
def projection_ops(q):
    """L_4f1 (4q^2 x q^2): a face panel's Gauss grid to its four quadrants'
    (A, B, C, D = (lo, lo), (hi, lo), (hi, hi), (lo, hi) in the two face
    coordinates); L_1f4 (q^2 x 4q^2): back, each coarse point interpolated
    from the quadrant that holds it."""
    if q % 2:
        raise ValueError("q must be even")
    g = osp.gauss_nodes(q)
    lo, hi = 0.5 * g - 0.5, 0.5 * g + 0.5
    quads = ((lo, lo), (hi, lo), (hi, hi), (lo, hi))
    L41 = np.concatenate([np.kron(osp.lagrange_matrix(g, a), osp.lagrange_matrix(g, b)) for a, b in quads])
    L14 = np.zeros((q * q, 4 * q * q))
    h = q // 2
    for k, (a, b) in enumerate(quads):
        ia = np.arange(q) < h if a is lo else np.arange(q) >= h
        ib = np.arange(q) < h if b is lo else np.arange(q) >= h
        rows = np.nonzero(np.outer(ia, ib).ravel())[0]
        M = np.kron(osp.lagrange_matrix(a, g[ia]), osp.lagrange_matrix(b, g[ib]))
        L14[rows, k * q * q:(k + 1) * q * q] = M
    return L41, L14


# ---------------------------------------------------------------------------
# compression lists (merge/_utils_adaptive_3D_DtN.py:11-166)
# ---------------------------------------------------------------------------
def _walk(l0, l1):
    o0, o1 = [], []
    i = j = 0
    while i < len(l0) and j < len(l1):
        s0, s1 = side(l0[i]), side(l1[j])
        if s0 == s1:
            o0.append(False)
            o1.append(False)
            i += 1
            j += 1
        elif s0 < s1:
            o0 += [True] * 4
            o1.append(False)
            i += 4
            j += 1
        else:
            o1 += [True] * 4
            o0.append(False)
            j += 4
            i += 1
    return o0, o1


def compression_lists(nd):
    """For interfaces 9..20 (IFACES order): (flags of the first child's face,
    flags of the second's)."""
    ch = nd.children
    return [_walk(face_leaves(ch[c1], f1), face_leaves(ch[c2], f2)) for (c1, f1), (c2, f2) in IFACES]


# ---------------------------------------------------------------------------
# leaf stage
# ---------------------------------------------------------------------------
def leaf_stage(root, p, q, coeffs, source):
    lv = leaves(root)
    f = source if source.ndim == 3 else source[..., None]
    cache = {}
    order = ("D_xx", "D_xy", "D_yy", "D_xz", "D_yz", "D_zz", "D_x", "D_y", "D_z", "I")
    Ys, Ts, vs, hs = [], [], [], []
    for i, lf in enumerate(lv):
        half = side(lf) / 2
        if half not in cache:
            cache[half] = hps3d.leaf_ops_3D(p, q, half)
        ops = cache[half]
        A = hps2d.assemble_diff_operator({k: v[i:i + 1] for k, v in coeffs.items()}, ops, order=order)
        Y, T, v, h = hps2d.leaf_dtn(A, f[i:i + 1], ops["P"], ops["Q"])
        Ys.append(Y[0])
        Ts.append(T[0])
        vs.append(v[0])
        hs.append(h[0])
    return np.array(Ys), np.array(Ts), np.array(vs), np.array(hs)


# ---------------------------------------------------------------------------
# merges
# ---------------------------------------------------------------------------
def _compress_rows(M, L14, flags, q2):
    out, i = [], 0
    while i < len(flags):
        if flags[i]:
            out.append(L14 @ M[q2 * i:q2 * (i + 4)])
            i += 4
        else:
            out.append(M[q2 * i:q2 * (i + 1)])
            i += 1
    return np.concatenate(out, axis=0)


def _compressed_child(T, h, counts, flags, q, L41, L14):
    """Child map with the flagged panels of each face compressed; returns
    (T', h', offsets, lengths) of the faces in the compressed order."""
    q2 = q * q
    off = np.concatenate([[0], np.cumsum(counts)])
    idx = [np.arange(off[f], off[f + 1]) for f in range(6)]
    rows_T, rows_h = [], []
    for f in range(6):
        rows_T.append(_compress_rows(T[idx[f]], L14, flags[f], q2))
        rows_h.append(_compress_rows(h[idx[f]], L14, flags[f], q2))
    Tr = np.concatenate(rows_T, axis=0)
    cols = [_compress_rows(Tr[:, idx[f]].T, L41.T, flags[f], q2).T for f in range(6)]
    Tc = np.concatenate(cols, axis=1)
    lens = [r.shape[0] for r in rows_T]
    return Tc, np.concatenate(rows_h, axis=0), np.concatenate([[0], np.cumsum(lens)]), lens


def oct_merge(nd, Ts, hs, q):
    """One 8:1 merge of nd's children: S, T, h_out, g_tilde."""
    L41, L14 = projection_ops(q)
    q2 = q * q
    fl = compression_lists(nd)
    flags = [[[False] * (face_counts(nd.children[c], q)[f] // q2) for f in range(6)] for c in range(8)]
    for i, ((c1, f1), (c2, f2)) in enumerate(IFACES):
        flags[c1][f1], flags[c2][f2] = list(fl[i][0]), list(fl[i][1])
    comp = [_compressed_child(Ts[c], hs[c], face_counts(nd.children[c], q), flags[c], q, L41, L14) for c in range(8)]

    def blk(c, f):
        o, L = comp[c][2], comp[c][3]
        return slice(int(o[f]), int(o[f]) + L[f])

    ext = [(c, k) for k, kids in enumerate(EXT_FACES) for c in kids]
    isz = [comp[c1][3][f1] for (c1, f1), _ in IFACES]
    assert all(comp[c2][3][f2] == isz[i] for i, (_, (c2, f2)) in enumerate(IFACES))
    esz = [comp[c][3][f] for c, f in ext]
    io = np.concatenate([[0], np.cumsum(isz)])
    eo = np.concatenate([[0], np.cumsum(esz)])
    ni, ne = io[-1], eo[-1]
    nsrc = hs[0].shape[-1]
    dt = np.result_type(Ts[0].dtype, hs[0].dtype)
    D = np.zeros((ni, ni), dtype=dt)
    C = np.zeros((ni, ne), dtype=dt)
    B = np.zeros((ne, ni), dtype=dt)
    A = np.zeros((ne, ne), dtype=dt)
    h_int = np.zeros((ni, nsrc), dtype=dt)
    h_ext = np.zeros((ne, nsrc), dtype=dt)
    I = lambda i: slice(int(io[i]), int(io[i + 1]))
    E = lambda e: slice(int(eo[e]), int(eo[e + 1]))
    for i, sides_i in enumerate(IFACES):
        for c, f in sides_i:
            Tc, hc = comp[c][0], comp[c][1]
            h_int[I(i)] += hc[blk(c, f)]
            for j, sides_j in enumerate(IFACES):
                for c2, f2 in sides_j:
                    if c2 == c:
                        D[I(i), I(j)] += Tc[blk(c, f), blk(c, f2)]
            for e, (ce, fe) in enumerate(ext):
                if ce == c:
                    C[I(i), E(e)] += Tc[blk(c, f), blk(c, fe)]
    for e, (ce, fe) in enumerate(ext):
        Tc, hc = comp[ce][0], comp[ce][1]
        h_ext[E(e)] = hc[blk(ce, fe)]
        for e2, (ce2, fe2) in enumerate(ext):
            if ce2 == ce:
                A[E(e), E(e2)] = Tc[blk(ce, fe), blk(ce, fe2)]
        for j, sides_j in enumerate(IFACES):
            for c2, f2 in sides_j:
                if c2 == ce:
                    B[E(e), I(j)] += Tc[blk(ce, fe), blk(ce, f2)]
    Dinv = np.linalg.inv(D)
    S = -Dinv @ C
    gt = -Dinv @ h_int
    return S, A + B @ S, h_ext + B @ gt, gt, (comp, flags, io, eo)


def merge_stage(root, T_leaf, h_leaf, q):
    Tm, hm = {}, {}
    for lf, T, h in zip(leaves(root), T_leaf, h_leaf):
        Tm[id(lf)], hm[id(lf)] = T, h
    info = {}

    def rec(nd):
        if not nd.children:
            return
        for c in nd.children:
            rec(c)
        S, T, h, gt, aux = oct_merge(nd, [Tm[id(c)] for c in nd.children], [hm[id(c)] for c in nd.children], q)
        Tm[id(nd)], hm[id(nd)], info[id(nd)] = T, h, (S, gt, aux)

    rec(root)
    return info, Tm[id(root)]


def down_pass(root, bdry, info, Y, v, q):
    L41, _ = projection_ops(q)
    q2 = q * q
    g_leaf = {}

    def decompress(x, flags):
        out, i, k = [], 0, 0
        while i < len(flags):
            panel = x[k:k + q2]
            if flags[i]:
                out.append(L41 @ panel)
                i += 4
            else:
                out.append(panel)
                i += 1
            k += q2
        return np.concatenate(out)

    def rec(nd, g):
        if not nd.children:
            g_leaf[id(nd)] = g
            return
        S, gt, (comp, flags, io, eo) = info[id(nd)]
        gi = S @ g + gt
        ext = [(c, k) for k, kids in enumerate(EXT_FACES) for c in kids]
        faces = [[None] * 6 for _ in range(8)]
        for e, (c, f) in enumerate(ext):
            faces[c][f] = g[int(eo[e]):int(eo[e + 1])]
        for i, sides in enumerate(IFACES):
            for c, f in sides:
                faces[c][f] = decompress(gi[int(io[i]):int(io[i + 1])], flags[c][f])
        for c, ch in enumerate(nd.children):
            rec(ch, np.concatenate(faces[c]))

    rec(root, bdry)
    return np.array([Y[i] @ g_leaf[id(lf)] + v[i] for i, lf in enumerate(leaves(root))])


def solve(root, p, q, coeffs, source, bdry):
    Y, T, v, h = leaf_stage(root, p, q, coeffs, source)
    info, Troot = merge_stage(root, T, h, q)
    b = bdry if bdry.ndim == 2 else bdry[:, None]
    return down_pass(root, b, info, Y, v, q), Troot
