"""ORACLE (test infrastructure only) -- 3D uniform DtN build/solve restated
on the CPU in FP64.

Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` may import this module; it is the checker, never the product.

* Leaf operators from coordinates: the p^3 grid, its boundary-first order
  (faces x-, x+, y-, y+, z-, z+ then the interior, each in ascending
  meshgrid(ij) index -- ``_grid_creation_3D.py:421-458``), D matrices by
  Kronecker products of the barycentric 1D matrix, P by averaging the 2D
  Lagrange interpolants of every face a boundary point lies on
  (``_precompute_operators_3D.py:64-147`` does this with index lists), Q as
  the interpolated outward normal derivative (:150-210).
* Oct merges from the interface/face description of
  ``merge/_uniform_3D_DtN.py:237-541`` and ``merge/_schur_complement.py:294-774``.
* The global DtN system assembled from coordinates and solved densely
  (pins every ordering convention), and manufactured polynomial solutions
  that the discretisation reproduces exactly.
"""

from __future__ import annotations

import numpy as np

from .spectral import cheb_points, diff_matrix, gauss_nodes, lagrange_matrix


# ---------------------------------------------------------------------------
# leaf
# ---------------------------------------------------------------------------
def leaf_order_3D(p: int):
    """(perm, XYZ) of the boundary-first leaf order on [-1, 1]^3."""
    x = cheb_points(p)
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    pts = np.stack([X.ravel(), Y.ravel(), Z.ravel()], -1)
    lo, hi = x[0], x[-1]
    taken = np.zeros(p**3, dtype=bool)
    order = []
    for axis in range(3):
        for val in (lo, hi):
            sel = np.nonzero((pts[:, axis] == val) & ~taken)[0]
            order.extend(sel.tolist())
            taken[sel] = True
    order.extend(np.nonzero(~taken)[0].tolist())
    perm = np.array(order)
    return perm, pts[perm]


# (fixed axis, value sign, first face coordinate, second face coordinate) of faces 1..6
FACES = ((0, -1, 1, 2), (0, 1, 1, 2), (1, -1, 0, 2), (1, 1, 0, 2), (2, -1, 0, 1), (2, 1, 0, 1))


def leaf_ops_3D(p: int, q: int, half_side: float):
    """D_x .. D_yz (p^3 x p^3), P (p^3-(p-2)^3 x 6q^2), Q (6q^2 x p^3) on one leaf."""
    perm, pts = leaf_order_3D(p)
    x = cheb_points(p)
    D1 = diff_matrix(x) / half_side
    I = np.eye(p)
    Dn = [np.kron(D1, np.kron(I, I)), np.kron(I, np.kron(D1, I)), np.kron(I, np.kron(I, D1))]
    Dx, Dy, Dz = (D[perm][:, perm] for D in Dn)
    ops = dict(D_x=Dx, D_y=Dy, D_z=Dz, D_xx=Dx @ Dx, D_yy=Dy @ Dy, D_zz=Dz @ Dz, D_xy=Dx @ Dy, D_xz=Dx @ Dz,
               D_yz=Dy @ Dz)
    nb = p**3 - (p - 2) ** 3
    g = gauss_nodes(q)
    P = np.zeros((nb, 6 * q * q))
    for r in range(nb):
        on = [k for k, (ax, sg, _, _) in enumerate(FACES) if pts[r, ax] == sg * 1.0]
        assert on, r
        for k in on:
            _, _, a1, a2 = FACES[k]
            la = lagrange_matrix(g, np.array([pts[r, a1]]))[0]
            lb = lagrange_matrix(g, np.array([pts[r, a2]]))[0]
            P[r, k * q * q : (k + 1) * q * q] += np.outer(la, lb).ravel() / len(on)
    Q = np.zeros((6 * q * q, p**3))
    Dax = (Dx, Dy, Dz)
    for k, (ax, sg, a1, a2) in enumerate(FACES):
        face = np.nonzero(pts[:, ax] == sg * 1.0)[0]
        # Lagrange weights from the face's Chebyshev points to each Gauss point
        la = lagrange_matrix(x, g)  # (q, p): l_i(g_a)
        for ia in range(q):
            for ib in range(q):
                row = k * q * q + ia * q + ib
                for c in face:
                    ci = int(np.argmin(np.abs(x - pts[c, a1])))
                    cj = int(np.argmin(np.abs(x - pts[c, a2])))
                    w = la[ia, ci] * la[ib, cj]
                    if w != 0.0:
                        Q[row] += sg * w * Dax[ax][c]
    ops["P"], ops["Q"] = P, Q
    return ops


def oct_children(b):
    """Children a..h of [x0, x1, y0, y1, z0, z1] (``_grid_creation_3D.py:29-78``)."""
    x0, x1, y0, y1, z0, z1 = b
    xm, ym, zm = (x0 + x1) / 2, (y0 + y1) / 2, (z0 + z1) / 2
    quads = [(x0, xm, y0, ym), (xm, x1, y0, ym), (xm, x1, ym, y1), (x0, xm, ym, y1)]
    return [qd + (zm, z1) for qd in quads] + [qd + (z0, zm) for qd in quads]


def uniform_leaves_3D(bounds, L):
    boxes = [tuple(bounds)]
    for _ in range(L):
        boxes = [c for b in boxes for c in oct_children(b)]
    return boxes


def leaf_points_3D(bounds, L, p):
    _, pts = leaf_order_3D(p)
    out = []
    for b in uniform_leaves_3D(bounds, L):
        c = np.array([(b[0] + b[1]) / 2, (b[2] + b[3]) / 2, (b[4] + b[5]) / 2])
        h = np.array([(b[1] - b[0]) / 2, (b[3] - b[2]) / 2, (b[5] - b[4]) / 2])
        out.append(c + h * pts)
    return np.stack(out)


def _face_patch_points(lo1, hi1, lo2, hi2, q):
    g = gauss_nodes(q)
    a = (lo1 + hi1) / 2 + (hi1 - lo1) / 2 * g
    b = (lo2 + hi2) / 2 + (hi2 - lo2) / 2 * g
    A, B = np.meshgrid(a, b, indexing="ij")
    return np.stack([A.ravel(), B.ravel()], -1)


def leaf_gauss_points_3D(b, q):
    """A leaf's 6 q^2 Gauss boundary points, faces 1..6, each q x q in (first, second) ij order."""
    lims = [(b[0], b[1]), (b[2], b[3]), (b[4], b[5])]
    out = []
    for ax, sg, a1, a2 in FACES:
        fp = _face_patch_points(*lims[a1], *lims[a2], q)
        xyz = np.zeros((q * q, 3))
        xyz[:, ax] = lims[ax][1] if sg > 0 else lims[ax][0]
        xyz[:, a1] = fp[:, 0]
        xyz[:, a2] = fp[:, 1]
        out.append(xyz)
    return np.concatenate(out)


def _quad_patches(lo1, hi1, lo2, hi2, L):
    boxes = [(lo1, hi1, lo2, hi2)]
    for _ in range(L):
        nxt = []
        for (a0, a1, b0, b1) in boxes:
            am, bm = (a0 + a1) / 2, (b0 + b1) / 2
            nxt += [(a0, am, b0, bm), (am, a1, b0, bm), (am, a1, bm, b1), (a0, am, bm, b1)]
        boxes = nxt
    return boxes


def boundary_gauss_points_3D(bounds, L, q):
    """Root boundary Gauss points: faces 1..6, each face's patches in recursive
    quad order of its two coordinates (``_grid_creation_3D.py:119-190``)."""
    lims = [(bounds[0], bounds[1]), (bounds[2], bounds[3]), (bounds[4], bounds[5])]
    out = []
    for ax, sg, a1, a2 in FACES:
        for patch in _quad_patches(*lims[a1], *lims[a2], L):
            fp = _face_patch_points(*patch, q)
            xyz = np.zeros((q * q, 3))
            xyz[:, ax] = lims[ax][1] if sg > 0 else lims[ax][0]
            xyz[:, a1] = fp[:, 0]
            xyz[:, a2] = fp[:, 1]
            out.append(xyz)
    return np.concatenate(out)


# ---------------------------------------------------------------------------
# oct merge
# ---------------------------------------------------------------------------
# interfaces 9..20: (child, face) pairs; faces 0..5 = x-, x+, y-, y+, z-, z+
IFACES = [((0, 1), (1, 0)), ((1, 3), (2, 2)), ((2, 0), (3, 1)), ((3, 2), (0, 3)),
          ((4, 1), (5, 0)), ((5, 3), (6, 2)), ((6, 0), (7, 1)), ((7, 2), (4, 3)),
          ((0, 4), (4, 5)), ((1, 4), (5, 5)), ((2, 4), (6, 5)), ((3, 4), (7, 5))]
# merged faces 1..6 = the listed children's same-numbered face blocks
EXT_FACES = [(4, 7, 3, 0), (5, 6, 2, 1), (4, 5, 1, 0), (7, 6, 2, 3), (4, 5, 6, 7), (0, 1, 2, 3)]


def oct_merge_dtn(T8, h8):
    """One 8:1 DtN merge: T8 (8, 6n, 6n), h8 (8, 6n, nsrc) -> S (12n, 24n), T (24n, 24n),
    h_ext (24n, nsrc), g_tilde (12n, nsrc) in the reference's orders."""
    n = T8.shape[1] // 6

    def blk(f):
        return slice(f * n, (f + 1) * n)

    ext = [(c, k) for k, kids in enumerate(EXT_FACES) for c in kids]
    ni, ne = 12 * n, 24 * n
    dt = np.result_type(T8.dtype, h8.dtype)
    D = np.zeros((ni, ni), dtype=dt)
    C = np.zeros((ni, ne), dtype=dt)
    B = np.zeros((ne, ni), dtype=dt)
    A = np.zeros((ne, ne), dtype=dt)
    h_int = np.zeros((ni, h8.shape[-1]), dtype=dt)
    h_ext = np.zeros((ne, h8.shape[-1]), dtype=dt)
    for i, sides_i in enumerate(IFACES):
        for c, f in sides_i:
            h_int[blk(i)] += h8[c][blk(f)]
            for j, sides_j in enumerate(IFACES):
                for c2, f2 in sides_j:
                    if c2 == c:
                        D[blk(i), blk(j)] += T8[c][blk(f), blk(f2)]
            for e, (ce, fe) in enumerate(ext):
                if ce == c:
                    C[blk(i), blk(e)] += T8[c][blk(f), blk(fe)]
    for e, (ce, fe) in enumerate(ext):
        h_ext[blk(e)] = h8[ce][blk(fe)]
        for e2, (ce2, fe2) in enumerate(ext):
            if ce2 == ce:
                A[blk(e), blk(e2)] = T8[ce][blk(fe), blk(fe2)]
        for j, sides_j in enumerate(IFACES):
            for c2, f2 in sides_j:
                if c2 == ce:
                    B[blk(e), blk(j)] += T8[ce][blk(fe), blk(f2)]
    Dinv = np.linalg.inv(D)
    S = -Dinv @ C
    gt = -Dinv @ h_int
    return S, A + B @ S, h_ext + B @ gt, gt


def merge_stage_3d(T, h, L):
    """L levels of oct merges; every level keeps a leading merge axis here."""
    S_lst, g_lst = [], []
    for _ in range(L):
        nm = T.shape[0] // 8
        outs = [oct_merge_dtn(T[8 * i : 8 * i + 8], h[8 * i : 8 * i + 8]) for i in range(nm)]
        S_lst.append(np.stack([o[0] for o in outs]))
        T = np.stack([o[1] for o in outs])
        h = np.stack([o[2] for o in outs])
        g_lst.append(np.stack([o[3] for o in outs]))
    return S_lst, g_lst, T[0], h[0]


def propagate_down_oct(S, g, gt):
    """Children's boundary data (8, 6n, nsrc) from the parent's g (24n, nsrc)."""
    n = g.shape[0] // 24
    gi = S @ g + gt
    out = np.zeros((8, 6 * n, g.shape[-1]), dtype=gi.dtype)
    for e, (c, f) in enumerate([(c, k) for k, kids in enumerate(EXT_FACES) for c in kids]):
        out[c, f * n : (f + 1) * n] = g[e * n : (e + 1) * n]
    for i, sides in enumerate(IFACES):
        for c, f in sides:
            out[c, f * n : (f + 1) * n] = gi[i * n : (i + 1) * n]
    return out


def down_pass_3d(bdry, S_lst, g_lst, Y, v):
    g = bdry[None]
    for S, gt in zip(reversed(S_lst), reversed(g_lst)):
        g = np.concatenate([propagate_down_oct(S[i], g[i], gt[i]) for i in range(S.shape[0])])
    return Y @ g + v, g


# ---------------------------------------------------------------------------
# global system
# ---------------------------------------------------------------------------
def _key3(pt):
    return (round(float(pt[0]), 11), round(float(pt[1]), 11), round(float(pt[2]), 11))


def solve_dtn_global_3d(A, f, P, Q, bounds, L, q, g_root, root_pts):
    """Dense solve of the 3D DtN HPS system: leaf rows A_ii u_i + A_ie P g_leaf = f_i,
    flux balance on every interior Gauss node, Dirichlet data on the root."""
    nl, n, _ = A.shape
    nb = P.shape[0]
    ni = n - nb
    root_val = {_key3(pt): val for pt, val in zip(root_pts, g_root)}
    leaf_nodes = [[_key3(pt) for pt in leaf_gauss_points_3D(b, q)] for b in uniform_leaves_3D(bounds, L)]
    interior = sorted({k for ln in leaf_nodes for k in ln if k not in root_val})
    gid = {k: i for i, k in enumerate(interior)}
    N = nl * ni + len(interior)
    M = np.zeros((N, N))
    rhs = np.zeros(N)
    touch = {k: [] for k in interior}
    for l, ln in enumerate(leaf_nodes):
        for a, k in enumerate(ln):
            if k in touch:
                touch[k].append((l, a))
    for l in range(nl):
        r0 = l * ni
        M[r0 : r0 + ni, r0 : r0 + ni] = A[l, nb:, nb:]
        AieP = A[l, nb:, :nb] @ P
        rhs[r0 : r0 + ni] = f[l, nb:]
        for a, k in enumerate(leaf_nodes[l]):
            if k in gid:
                M[r0 : r0 + ni, nl * ni + gid[k]] += AieP[:, a]
            else:
                rhs[r0 : r0 + ni] -= AieP[:, a] * root_val[k]
    for k, j in gid.items():
        row = nl * ni + j
        assert len(touch[k]) == 2, k
        for (l, a) in touch[k]:
            Qb = Q[a, :nb] @ P
            M[row, l * ni : (l + 1) * ni] += Q[a, nb:]
            for b, kk in enumerate(leaf_nodes[l]):
                if kk in gid:
                    M[row, nl * ni + gid[kk]] += Qb[b]
                else:
                    rhs[row] -= Qb[b] * root_val[kk]
    x = np.linalg.solve(M, rhs)
    u = np.zeros((nl, n))
    for l in range(nl):
        gl = np.array([x[nl * ni + gid[k]] if k in gid else root_val[k] for k in leaf_nodes[l]])
        u[l, :nb] = P @ gl
        u[l, nb:] = x[l * ni : (l + 1) * ni]
    return u
