"""ORACLE (test infrastructure only) -- the uniform HPS discrete system as one
sparse matrix, solved by a sparse direct factorisation with iterative
refinement in extended precision.

Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` may import this package.

The uniform build+solve of the reference eliminates, exactly, one global
linear system: every leaf's collocation rows plus the interface coupling on
the Gauss nodes.  ``global_system`` writes the same system densely for tiny
trees; this module assembles it sparsely (scipy.sparse), matching Gauss
nodes between leaves by their coordinates (not by the reference's index
tables), so trees up to L = 5 at p = 16 fit.  Its solution is the exact
discrete solution to within the extended-precision refinement's accuracy,
which makes it the yardstick that tells whether a stage-by-stage FP64 solve
(the oracle ``hps2d``, the GPU paths) or a refined mixed-precision solve is
closer to the discrete solution.

ItI system (``merge/_uniform_2D_ItI.py:183-375``, leaf rows of
``local_solve/_uniform_2D_ItI.py:120-193``), unknowns per leaf l: u_l (n
values) then its incoming impedance t_l (4q values):
  G u_l - P t_l = 0,   A_l[nb:, :] u_l = f_l[nb:]
  t_l(node) + (QH u_l')(node') = 0   for the neighbour l' across an interface
  t_l(node) = given incoming impedance on the root boundary.
DtN system (``merge/_schur_complement.py:183-237``), unknowns per leaf: its
interior values, then one Dirichlet value per interior Gauss node:
  A_ii u_i + A_ie P g_leaf = f_i;  the two leaves' fluxes (Q u)(node) sum to 0.
"""

from __future__ import annotations

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from .spectral import leaf_gauss_points_2D, uniform_leaves_2D


def _keys(pts):
    """Integer keys of points, for matching coordinates across leaves."""
    return [tuple(k) for k in np.round(np.asarray(pts) * 1e10).astype(np.int64)]


def _leaf_nodes(bounds, L, q):
    boxes = uniform_leaves_2D(bounds, L)
    return [_keys(leaf_gauss_points_2D(b, q)) for b in boxes]


def iti_sparse_system(A, f, P_iti, QH, G, bounds, L, q, t_root, root_pts):
    """(K csr, rhs (N, nsrc), per) of the ItI system; f (nl, n, nsrc),
    t_root (nroot, nsrc)."""
    nl, n, _ = A.shape
    nb = P_iti.shape[0]
    nq = 4 * q
    per = n + nq
    N = nl * per
    nsrc = f.shape[-1]
    leaf_nodes = _leaf_nodes(bounds, L, q)
    root_idx = {k: i for i, k in enumerate(_keys(root_pts))}
    owner = {}
    for l, ln in enumerate(leaf_nodes):
        for a, k in enumerate(ln):
            owner.setdefault(k, []).append((l, a))
    rows, cols, vals = [], [], []
    rhs = np.zeros((N, nsrc), dtype=np.complex128)
    gi, gj = np.nonzero(G)
    pi, pj = np.nonzero(P_iti)
    ai, aj = np.meshgrid(np.arange(nb, n), np.arange(n), indexing="ij")
    for l in range(nl):
        u0 = l * per
        t0 = u0 + n
        rows += [u0 + gi, u0 + pi, u0 + ai.ravel()]
        cols += [u0 + gj, t0 + pj, u0 + aj.ravel()]
        vals += [G[gi, gj], -P_iti[pi, pj], A[l, nb:, :].ravel()]
        rhs[u0 + nb : u0 + n] = f[l, nb:]
        trow = []
        tcol = []
        tval = []
        for a, k in enumerate(leaf_nodes[l]):
            row = t0 + a
            trow.append(row)
            tcol.append(row)
            tval.append(1.0)
            if k in root_idx:
                rhs[row] = t_root[root_idx[k]]
            else:
                (lb, b), = [(lb, b) for (lb, b) in owner[k] if lb != l]
                trow += [row] * n
                tcol += list(range(lb * per, lb * per + n))
                tval += list(QH[b])
        rows.append(np.asarray(trow))
        cols.append(np.asarray(tcol))
        vals.append(np.asarray(tval, dtype=np.complex128))
    K = sp.csr_matrix((np.concatenate(vals).astype(np.complex128), (np.concatenate(rows), np.concatenate(cols))),
                      shape=(N, N))
    return K, rhs, per


def dtn_sparse_system(A, f, P, Q, bounds, L, q, g_root, root_pts):
    """(K csr, rhs (N, nsrc), layout) of the DtN system; layout = (ni,
    {node key: global index}, leaf node keys, root index map)."""
    nl, n, _ = A.shape
    nb = P.shape[0]
    ni = n - nb
    nsrc = f.shape[-1]
    leaf_nodes = _leaf_nodes(bounds, L, q)
    root_idx = {k: i for i, k in enumerate(_keys(root_pts))}
    interior = sorted({k for ln in leaf_nodes for k in ln if k not in root_idx})
    gid = {k: nl * ni + i for i, k in enumerate(interior)}
    N = nl * ni + len(interior)
    rows, cols, vals = [], [], []
    rhs = np.zeros((N, nsrc), dtype=np.result_type(A.dtype, f.dtype, g_root.dtype))
    ri, rj = np.meshgrid(np.arange(ni), np.arange(ni), indexing="ij")
    QbP = Q[:, :nb] @ P
    for l in range(nl):
        r0 = l * ni
        rows.append(r0 + ri.ravel())
        cols.append(r0 + rj.ravel())
        vals.append(A[l, nb:, nb:].ravel())
        rhs[r0 : r0 + ni] += f[l, nb:]
        AieP = A[l, nb:, :nb] @ P
        for a, k in enumerate(leaf_nodes[l]):
            if k in gid:
                rows.append(r0 + np.arange(ni))
                cols.append(np.full(ni, gid[k]))
                vals.append(AieP[:, a])
            else:
                rhs[r0 : r0 + ni] -= AieP[:, a, None] * g_root[root_idx[k]][None, :]
            if k in gid:
                row = gid[k]
                rows.append(np.full(ni, row))
                cols.append(r0 + np.arange(ni))
                vals.append(Q[a, nb:])
                for b, kk in enumerate(leaf_nodes[l]):
                    if kk in gid:
                        rows.append(np.array([row]))
                        cols.append(np.array([gid[kk]]))
                        vals.append(np.array([QbP[a, b]]))
                    else:
                        rhs[row] -= QbP[a, b] * g_root[root_idx[kk]]
    K = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(N, N))
    return K, rhs, (ni, gid, leaf_nodes, root_idx)


def _matvec_ld(K, x):
    """K @ x with products and sums in extended precision (x (N, nsrc))."""
    Kc = K.tocsr()
    xl = x.astype(np.clongdouble if np.iscomplexobj(x) or np.iscomplexobj(Kc.data) else np.longdouble)
    dl = Kc.data.astype(np.clongdouble if np.iscomplexobj(Kc.data) else np.longdouble)
    out = np.zeros(x.shape, dtype=xl.dtype)
    counts = np.diff(Kc.indptr)
    nz = counts > 0
    prod = dl[:, None] * xl[Kc.indices]
    sums = np.add.reduceat(prod, Kc.indptr[:-1][nz], axis=0)
    out[nz] = sums
    return out


def solve_refined(K, rhs, iters=3):
    """x = K^{-1} rhs by SuperLU, then `iters` refinement steps whose residuals
    are formed in extended precision.  Returns (x, relative residual history)."""
    lu = spla.splu(K.tocsc(), permc_spec="COLAMD")
    x = lu.solve(rhs)
    rl = rhs.astype(np.clongdouble if np.iscomplexobj(rhs) or np.iscomplexobj(K.data) else np.longdouble)
    hist = []
    for _ in range(iters):
        r = rl - _matvec_ld(K, x)
        hist.append(float(np.abs(r).max() / max(np.abs(rl).max(), 1e-300)))
        x = x + lu.solve(r.astype(x.dtype))
    return x, hist


def solve_iti_sparse(A, f, P_iti, QH, G, bounds, L, q, t_root, root_pts, iters=3):
    """Exact discrete ItI solution u (nl, n, nsrc)."""
    K, rhs, per = iti_sparse_system(A, f, P_iti, QH, G, bounds, L, q, t_root, root_pts)
    x, _ = solve_refined(K, rhs, iters)
    nl, n, _ = A.shape
    return np.stack([x[l * per : l * per + n] for l in range(nl)])


def solve_dtn_sparse(A, f, P, Q, bounds, L, q, g_root, root_pts, iters=3):
    """Exact discrete DtN solution u (nl, n, nsrc)."""
    K, rhs, (ni, gid, leaf_nodes, root_idx) = dtn_sparse_system(A, f, P, Q, bounds, L, q, g_root, root_pts)
    x, _ = solve_refined(K, rhs, iters)
    nl, n, _ = A.shape
    nb = P.shape[0]
    u = np.zeros((nl, n, rhs.shape[1]), dtype=x.dtype)
    for l in range(nl):
        g = np.stack([x[gid[k]] if k in gid else g_root[root_idx[k]] for k in leaf_nodes[l]])
        u[l, :nb] = P @ g
        u[l, nb:] = x[l * ni : (l + 1) * ni]
    return u
