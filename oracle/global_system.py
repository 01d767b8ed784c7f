"""ORACLE (test infrastructure only) -- the HPS discrete system assembled
globally from coordinates and solved densely.

The uniform HPS build+solve of the reference computes, in exact arithmetic,
the solution of one global linear system (leaf collocation rows plus
interface coupling on the Gauss nodes).  This module writes that system down
directly -- matching Gauss nodes between neighbouring leaves by their
coordinates, not by the reference's index gathers, rolls and flips -- and
solves it with a dense LU.  Agreement with the stage-by-stage restatement in
``hps2d`` therefore pins every ordering convention of the merges and the down
pass.  Tiny inputs only (L <= 2, p <= 10).

DtN system (``merge/_schur_complement.py:183-237`` eliminates exactly this):
  leaf rows      A_ii u_i + A_ie P g_leaf = f_i
  interface rows sum over the two leaves of (Q [P g_leaf; u_i])(node) = 0
  root boundary  g = given Dirichlet data
ItI system (``merge/_uniform_2D_ItI.py:183-375``):
  leaf rows      G u - P t_in = 0,   A[nb:, :] u = f_int
  interface rows t_in,a(node) + (QH u_b)(node) = 0 for each side of an interface
  root boundary  t_in = given incoming impedance
"""

from __future__ import annotations

import numpy as np

from .spectral import leaf_gauss_points_2D, uniform_leaves_2D


def _key(pt):
    return (round(float(pt[0]), 11), round(float(pt[1]), 11))


def solve_dtn_global(A, f, P, Q, bounds, L, q, g_root, root_pts):
    """A: (nl, n, n) leaf operators, f: (nl, n), g_root: Dirichlet values at
    root_pts (root boundary Gauss points).  Returns u (nl, n)."""
    nl, n, _ = A.shape
    nb = P.shape[0]
    ni = n - nb
    boxes = uniform_leaves_2D(bounds, L)
    root_val = {_key(pt): val for pt, val in zip(root_pts, g_root)}
    leaf_nodes = [[_key(pt) for pt in leaf_gauss_points_2D(b, q)] for b in boxes]
    interior_nodes = sorted({k for ln in leaf_nodes for k in ln if k not in root_val})
    gid = {k: i for i, k in enumerate(interior_nodes)}
    N = nl * ni + len(interior_nodes)
    M = np.zeros((N, N), dtype=A.dtype)
    rhs = np.zeros(N, dtype=np.result_type(A.dtype, f.dtype, np.asarray(g_root).dtype))
    # which leaves touch each interior node, with local index
    touch = {k: [] for k in interior_nodes}
    for l, ln in enumerate(leaf_nodes):
        for a, k in enumerate(ln):
            if k in touch:
                touch[k].append((l, a))
    for l in range(nl):
        r0 = l * ni
        # A_ii u_i
        M[r0 : r0 + ni, r0 : r0 + ni] = A[l, nb:, nb:]
        AieP = A[l, nb:, :nb] @ P  # (ni, 4q)
        rhs[r0 : r0 + ni] = f[l, nb:]
        for a, k in enumerate(leaf_nodes[l]):
            if k in gid:
                M[r0 : r0 + ni, nl * ni + gid[k]] += AieP[:, a]
            else:
                rhs[r0 : r0 + ni] -= AieP[:, a] * root_val[k]
    for k, j in gid.items():
        row = nl * ni + j
        assert len(touch[k]) == 2
        for (l, a) in touch[k]:
            Qb = Q[a, :nb] @ P  # flux row, boundary part via P
            M[row, l * ni : (l + 1) * ni] += Q[a, nb:]
            for b, kk in enumerate(leaf_nodes[l]):
                if kk in gid:
                    M[row, nl * ni + gid[kk]] += Qb[b]
                else:
                    rhs[row] -= Qb[b] * root_val[kk]
    x = np.linalg.solve(M, rhs)
    u = np.zeros((nl, n), dtype=x.dtype)
    for l in range(nl):
        g = np.array([x[nl * ni + gid[k]] if k in gid else root_val[k] for k in leaf_nodes[l]])
        u[l, :nb] = P @ g
        u[l, nb:] = x[l * ni : (l + 1) * ni]
    return u


def solve_iti_global(A, f, P_iti, QH, G, bounds, L, q, t_root, root_pts):
    """ItI version; t_root is the incoming impedance at root_pts.  Returns u (nl, n)."""
    M, rhs, per = iti_system(A, f, P_iti, QH, G, bounds, L, q, t_root, root_pts)
    nl, n, _ = A.shape
    x = np.linalg.solve(M, rhs)
    return np.stack([x[l * per : l * per + n] for l in range(nl)])


def iti_system(A, f, P_iti, QH, G, bounds, L, q, t_root, root_pts):
    """The dense ItI system M x = rhs; unknowns per leaf l: u_l (n values) then
    its incoming impedance t_l (4q values), offset l * per.  Leaf l's interior
    rows are rows l*per + nb .. l*per + n, holding A[l, nb:, :] in columns
    l*per .. l*per + n."""
    nl, n, _ = A.shape
    nb = P_iti.shape[0]
    nq = 4 * q
    boxes = uniform_leaves_2D(bounds, L)
    root_val = {_key(pt): val for pt, val in zip(root_pts, t_root)}
    leaf_nodes = [[_key(pt) for pt in leaf_gauss_points_2D(b, q)] for b in boxes]
    owner = {}
    for l, ln in enumerate(leaf_nodes):
        for a, k in enumerate(ln):
            owner.setdefault(k, []).append((l, a))
    per = n + nq
    N = nl * per
    M = np.zeros((N, N), dtype=np.complex128)
    rhs = np.zeros(N, dtype=np.complex128)
    for l in range(nl):
        u0 = l * per
        t0 = u0 + n
        # G u - P t_in = 0
        M[u0 : u0 + nb, u0 : u0 + n] = G
        M[u0 : u0 + nb, t0 : t0 + nq] = -P_iti
        # A_int u = f_int
        M[u0 + nb : u0 + n, u0 : u0 + n] = A[l, nb:, :]
        rhs[u0 + nb : u0 + n] = f[l, nb:]
        for a, k in enumerate(leaf_nodes[l]):
            row = t0 + a
            M[row, t0 + a] = 1.0
            if k in root_val:
                rhs[row] = root_val[k]
            else:
                others = [(lb, b) for (lb, b) in owner[k] if lb != l]
                assert len(others) == 1
                lb, b = others[0]
                M[row, lb * per : lb * per + n] += QH[b]
    return M, rhs, per
