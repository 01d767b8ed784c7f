"""ORACLE (test infrastructure only) -- FP64 restatement of the reference's
uniform 2D HPS build/solve, stage by stage.

Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` may import this package.  Each function cites the reference
file:line it restates.  Arrays are numpy, leaf-batched exactly as in the
reference (row-major, leaf axis first).
"""

from __future__ import annotations

import numpy as np


# ---------------------------------------------------------------------------
# Leaf stage
# ---------------------------------------------------------------------------
def assemble_diff_operator(coeffs: dict, ops: dict, order=("D_xx", "D_xy", "D_yy", "D_x", "D_y", "I")):
    """A_leaf = sum_k diag(c_k) D_k over the coefficients that are present.

    ``local_solve/_uniform_2D_DtN.py:137-225`` (row scaling, fixed order
    [xx, xy, yy, x, y, I]).  coeffs maps name -> (n_leaves, n) array.
    """
    n = ops["D_xx"].shape[0]
    present = [k for k in order if coeffs.get(k) is not None]
    dtype = np.result_type(*[coeffs[k].dtype for k in present])
    nl = coeffs[present[0]].shape[0]
    A = np.zeros((nl, n, n), dtype=dtype)
    for k in present:
        D = np.eye(n) if k == "I" else ops[k]
        A += coeffs[k][:, :, None] * D[None, :, :]
    return A


def leaf_dtn(A: np.ndarray, source: np.ndarray, P: np.ndarray, Q: np.ndarray):
    """Batched ``get_DtN`` (``local_solve/_uniform_2D_DtN.py:228-273``).

    source: (n_leaves, n, nsrc).  Returns Y (nl, n, 4q), T (nl, 4q, 4q),
    v (nl, n, nsrc), h (nl, 4q, nsrc).
    """
    nb = P.shape[0]
    A_ii = A[:, nb:, nb:]
    A_ie = A[:, nb:, :nb]
    A_ii_inv = np.linalg.inv(A_ii)
    nl, n = A.shape[0], A.shape[1]
    L2 = np.zeros((nl, n, nb), dtype=A.dtype)
    L2[:, :nb] = np.eye(nb)
    L2[:, nb:] = -A_ii_inv @ A_ie
    Y = L2 @ P
    T = Q @ Y
    v = np.zeros((nl, n, source.shape[-1]), dtype=A.dtype)
    v[:, nb:] = A_ii_inv @ source[:, nb:]
    h = Q @ v
    return Y, T, v, h


def leaf_iti(A: np.ndarray, source: np.ndarray, P: np.ndarray, QH: np.ndarray, G: np.ndarray):
    """Batched ``get_ItI`` (``local_solve/_uniform_2D_ItI.py:120-193``).

    B = [G; A[nb:, :]] in complex128, Y = B^{-1}[:, :nb] P, T = QH Y,
    v = B^{-1}[:, nb:] f_int, h = QH v.  Returns (T, Y, h, v) like the reference.
    """
    nb = P.shape[0]
    nl, n = A.shape[0], A.shape[1]
    B = np.zeros((nl, n, n), dtype=np.complex128)
    B[:, :nb] = G
    B[:, nb:] = A[:, nb:]
    B_inv = np.linalg.inv(B)
    Phi = B_inv[:, :, nb:]
    Y = B_inv[:, :, :nb] @ P
    v = Phi @ source[:, nb:]
    h = QH @ v
    T = QH @ Y
    return T, Y, h, v


def leaf_iti_nosource(A, P, QH, G):
    """``get_ItI_nosource`` (``local_solve/_nosource_uniform_2D_ItI.py:81-146``):
    returns (T, Y, Phi) with Phi = B^{-1}[:, nb:]."""
    nb = P.shape[0]
    nl, n = A.shape[0], A.shape[1]
    B = np.zeros((nl, n, n), dtype=np.complex128)
    B[:, :nb] = G
    B[:, nb:] = A[:, nb:]
    B_inv = np.linalg.inv(B)
    Y = B_inv[:, :, :nb] @ P
    return QH @ Y, Y, B_inv[:, :, nb:]


def leaf_dtn_nosource(A, P, Q):
    """``get_DtN_nosource`` (``local_solve/_nosource_uniform_2D_DtN.py:77-123``):
    returns (T, Y, Phi) with Phi = A_ii^{-1}."""
    nb = P.shape[0]
    A_ii_inv = np.linalg.inv(A[:, nb:, nb:])
    nl, n = A.shape[0], A.shape[1]
    L2 = np.zeros((nl, n, nb), dtype=A.dtype)
    L2[:, :nb] = np.eye(nb)
    L2[:, nb:] = -A_ii_inv @ A[:, nb:, :nb]
    Y = L2 @ P
    return Q @ Y, Y, A_ii_inv


# ---------------------------------------------------------------------------
# Merge stage
# ---------------------------------------------------------------------------
def _blocks(T, h, i0, i1, i2):
    """``_get_submatrices`` (``merge/_uniform_2D_DtN.py:441-475``)."""
    ids = (i0, i1, i2)
    hs = [h[i] for i in ids]
    Ts = [[T[np.ix_(a, b)] for b in ids] for a in ids]
    return hs, Ts


def _quad_idx(n_ext):
    """Index sets of ``get_quadmerge_blocks_{a..d}`` (``merge/_uniform_2D_DtN.py:385-437``)."""
    m = n_ext // 4
    r = np.arange(n_ext)
    a = (np.concatenate([r[-m:], r[:m]]), r[m : 2 * m], r[2 * m : 3 * m][::-1])  # 1, 5, 8
    b = (r[: 2 * m], r[2 * m : 3 * m], r[3 * m :][::-1])  # 2, 6, 5
    c = (r[:m][::-1], r[m : 3 * m], r[3 * m :])  # 6, 3, 7
    d = (r[:m], r[m : 2 * m][::-1], r[2 * m :])  # 8, 7, 4
    return a, b, c, d


def quad_merge_dtn(Ta, Tb, Tc, Td, ha, hb, hc, hd):
    """One DtN 4:1 merge, ``_uniform_quad_merge_DtN`` (``merge/_uniform_2D_DtN.py:207-348``)
    with ``assemble_merge_outputs_DtN`` (``merge/_schur_complement.py:118-147,183-237``).
    h's are (n_ext, nsrc).  Returns S (4m, 8m), T (8m, 8m), h_out (8m, nsrc), g_tilde (4m, nsrc).
    """
    ia, ib, ic, id_ = _quad_idx(Ta.shape[0])
    (ha1, ha5, ha8), ((a11, a15, a18), (a51, a55, a58), (a81, a85, a88)) = _blocks(Ta, ha, *ia)
    (hb2, hb6, hb5), ((b22, b26, b25), (b62, b66, b65), (b52, b56, b55)) = _blocks(Tb, hb, *ib)
    (hc6, hc3, hc7), ((c66, c63, c67), (c36, c33, c37), (c76, c73, c77)) = _blocks(Tc, hc, *ic)
    (hd8, hd7, hd4), ((d88, d87, d84), (d78, d77, d74), (d48, d47, d44)) = _blocks(Td, hd, *id_)
    m, ne = a51.shape
    Zei = np.zeros((ne, m), dtype=Ta.dtype)
    Zie = np.zeros((m, ne), dtype=Ta.dtype)
    Zii = np.zeros((m, m), dtype=Ta.dtype)
    B = np.block([[a15, Zei, Zei, a18], [b25, b26, Zei, Zei], [Zei, c36, c37, Zei], [Zei, Zei, d47, d48]])
    C = np.block([[a51, b52, Zie, Zie], [Zie, b62, c63, Zie], [Zie, Zie, c73, d74], [a81, Zie, Zie, d84]])
    D = np.block(
        [
            [a55 + b55, b56, Zii, a58],
            [b65, b66 + c66, c67, Zii],
            [Zii, c76, c77 + d77, d78],
            [a85, Zii, d87, d88 + a88],
        ]
    )
    h_int = np.concatenate([ha5 + hb5, hb6 + hc6, hc7 + hd7, hd8 + ha8])
    h_ext = np.concatenate([ha1, hb2, hc3, hd4])
    D_inv = np.linalg.inv(D)
    S = -D_inv @ C
    T = B @ S
    o = 0
    for Ablk in (a11, b22, c33, d44):
        k = Ablk.shape[0]
        T[o : o + k, o : o + k] += Ablk
        o += k
    g_tilde = -D_inv @ h_int
    h_out = h_ext + B @ g_tilde
    h_out = np.roll(h_out, -m, axis=0)
    T = np.roll(np.roll(T, -m, axis=0), -m, axis=1)
    S = np.roll(S, -m, axis=1)
    return S, T, h_out, g_tilde


def _invert_D_iti(D12, D21):
    """``_invert_D_ItI`` (``merge/_schur_complement.py:79-114``)."""
    n = D12.shape[0]
    m = D21.shape[0]
    W = np.eye(n) - D12 @ D21
    Wi = np.linalg.inv(W)
    Di = np.zeros((n + m, n + m), dtype=D12.dtype)
    Di[:n, :n] = Wi
    Di[:n, n:] = -Wi @ D12
    Di[n:, :n] = -D21 @ Wi
    Di[n:, n:] = np.eye(m) + D21 @ Wi @ D12
    return Di


def _iti_blocks(Ra, Rb, Rc, Rd, ha, hb, hc, hd):
    ia, ib, ic, id_ = _quad_idx(Ra.shape[0])
    (ha1, ha5, ha8), ((a11, a15, a18), (a51, a55, a58), (a81, a85, a88)) = _blocks(Ra, ha, *ia)
    (hb2, hb6, hb5), ((b22, b26, b25), (b62, b66, b65), (b52, b56, b55)) = _blocks(Rb, hb, *ib)
    (hc6, hc3, hc7), ((c66, c63, c67), (c36, c33, c37), (c76, c73, c77)) = _blocks(Rc, hc, *ic)
    (hd8, hd7, hd4), ((d88, d87, d84), (d78, d77, d74), (d48, d47, d44)) = _blocks(Rd, hd, *id_)
    m, ne = a51.shape
    dt = Ra.dtype
    Zei = np.zeros((ne, m), dtype=dt)
    Zie = np.zeros((m, ne), dtype=dt)
    Zii = np.zeros((m, m), dtype=dt)
    B = np.block(
        [
            [a15, a18, Zei, Zei, Zei, Zei, Zei, Zei],
            [Zei, Zei, Zei, Zei, b25, b26, Zei, Zei],
            [Zei, Zei, c36, c37, Zei, Zei, Zei, Zei],
            [Zei, Zei, Zei, Zei, Zei, Zei, d47, d48],
        ]
    )
    C = np.block(
        [
            [Zie, b52, Zie, Zie],
            [Zie, Zie, Zie, d84],
            [Zie, b62, Zie, Zie],
            [Zie, Zie, Zie, d74],
            [a51, Zie, Zie, Zie],
            [Zie, Zie, c63, Zie],
            [Zie, Zie, c73, Zie],
            [a81, Zie, Zie, Zie],
        ]
    )
    D12 = np.block([[b55, b56, Zii, Zii], [Zii, Zii, d87, d88], [b65, b66, Zii, Zii], [Zii, Zii, d77, d78]])
    D21 = np.block([[a55, a58, Zii, Zii], [Zii, Zii, c66, c67], [Zii, Zii, c76, c77], [a85, a88, Zii, Zii]])
    h_int = np.concatenate([hb5, hd8, hb6, hd7, ha5, hc6, hc7, ha8])
    h_ext = np.concatenate([ha1, hb2, hc3, hd4])
    return m, (a11, b22, c33, d44), B, C, D12, D21, h_int, h_ext


def _iti_perm(m):
    """Row permutation r of ``merge/_uniform_2D_ItI.py:357-373``:
    [a5, a8, c6, c7, b5, b6, d7, d8] -> [a5, b5, b6, c6, c7, d7, d8, a8]."""
    return np.concatenate(
        [
            np.arange(m),
            np.arange(4 * m, 5 * m),
            np.arange(5 * m, 6 * m),
            np.arange(2 * m, 3 * m),
            np.arange(3 * m, 4 * m),
            np.arange(6 * m, 7 * m),
            np.arange(7 * m, 8 * m),
            np.arange(m, 2 * m),
        ]
    )


def quad_merge_iti(Ra, Rb, Rc, Rd, ha, hb, hc, hd):
    """One ItI 4:1 merge, ``_uniform_quad_merge_ItI`` (``merge/_uniform_2D_ItI.py:183-375``).
    Returns S (8m, 8m), T (8m, 8m), h_out (8m, nsrc), g_tilde (8m, nsrc)."""
    m, Alst, B, C, D12, D21, h_int, h_ext = _iti_blocks(Ra, Rb, Rc, Rd, ha, hb, hc, hd)
    D_inv = _invert_D_iti(D12, D21)
    S = -D_inv @ C
    T = B @ S
    o = 0
    for Ablk in Alst:
        k = Ablk.shape[0]
        T[o : o + k, o : o + k] += Ablk
        o += k
    g_tilde = -D_inv @ h_int
    h_out = h_ext + B @ g_tilde
    h_out = np.roll(h_out, -m, axis=0)
    T = np.roll(np.roll(T, -m, axis=0), -m, axis=1)
    S = np.roll(S, -m, axis=1)
    r = _iti_perm(m)
    return S[r], T, h_out, g_tilde[r]


def quad_merge_iti_nosource(Ra, Rb, Rc, Rd):
    """``_uniform_quad_merge_ItI`` of the no-source merge
    (``merge/_nosource_uniform_2D_ItI.py``): returns S (rows permuted by r),
    T, D^{-1} (unpermuted) and B D^{-1} (exterior rows in pre-roll order;
    the up pass rolls after applying it, ``up_pass/_uniform_2D_ItI.py:210-217``)."""
    ne = Ra.shape[0]
    z = np.zeros((ne, 1), dtype=Ra.dtype)
    m, Alst, B, C, D12, D21, _, _ = _iti_blocks(Ra, Rb, Rc, Rd, z, z, z, z)
    D_inv = _invert_D_iti(D12, D21)
    S = -D_inv @ C
    T = B @ S
    o = 0
    for Ablk in Alst:
        k = Ablk.shape[0]
        T[o : o + k, o : o + k] += Ablk
        o += k
    BD_inv = B @ D_inv  # exterior rows left in pre-roll order, as stored by the reference
    T = np.roll(np.roll(T, -m, axis=0), -m, axis=1)
    S = np.roll(S, -m, axis=1)
    r = _iti_perm(m)
    return S[r], T, D_inv, BD_inv


def merge_stage(T, h, L, kind="dtn"):
    """``merge_stage_uniform_2D_{DtN,ItI}`` (``merge/_uniform_2D_DtN.py:13-203``,
    ``merge/_uniform_2D_ItI.py:19-179``) for L levels of 4:1 merges.

    T: (n_leaves, ne, ne); h: (n_leaves, ne, nsrc).  Returns S_lst, g_tilde_lst
    (each level with a leading merge axis, the root with a leading axis of 1),
    and the top T (ne_root, ne_root), top h.
    """
    fn = quad_merge_dtn if kind == "dtn" else quad_merge_iti
    S_lst, g_lst = [], []
    for _ in range(L):
        nm = T.shape[0] // 4
        outs = [fn(T[4 * i], T[4 * i + 1], T[4 * i + 2], T[4 * i + 3], h[4 * i], h[4 * i + 1], h[4 * i + 2], h[4 * i + 3]) for i in range(nm)]
        S_lst.append(np.stack([o[0] for o in outs]))
        T = np.stack([o[1] for o in outs])
        h = np.stack([o[2] for o in outs])
        g_lst.append(np.stack([o[3] for o in outs]))
    return S_lst, g_lst, T[0], h[0]


def merge_stage_iti_nosource(T, L):
    """No-source ItI merge stage: S_lst, D_inv_lst, BD_inv_lst, top T."""
    S_lst, Di_lst, BDi_lst = [], [], []
    for _ in range(L):
        nm = T.shape[0] // 4
        outs = [quad_merge_iti_nosource(T[4 * i], T[4 * i + 1], T[4 * i + 2], T[4 * i + 3]) for i in range(nm)]
        S_lst.append(np.stack([o[0] for o in outs]))
        T = np.stack([o[1] for o in outs])
        Di_lst.append(np.stack([o[2] for o in outs]))
        BDi_lst.append(np.stack([o[3] for o in outs]))
    return S_lst, Di_lst, BDi_lst, T[0]


# ---------------------------------------------------------------------------
# Down pass
# ---------------------------------------------------------------------------
def propagate_down_dtn(S, g, gt):
    """``_propagate_down_2D_DtN`` (``down_pass/_uniform_2D_DtN.py:126-189``)."""
    n = g.shape[0] // 8
    gi = S @ g + gt
    g5, g6, g7, g8 = gi[:n], gi[n : 2 * n], gi[2 * n : 3 * n], gi[3 * n :]
    ga = np.concatenate([g[:n], g5, g8[::-1], g[7 * n :]])
    gb = np.concatenate([g[n : 3 * n], g6, g5[::-1]])
    gc = np.concatenate([g6[::-1], g[3 * n : 5 * n], g7])
    gd = np.concatenate([g8, g7[::-1], g[5 * n : 7 * n]])
    return np.stack([ga, gb, gc, gd])


def propagate_down_iti(S, g, gt):
    """``_propogate_down_2D_ItI`` (``down_pass/_uniform_2D_ItI.py:122-192``)."""
    n = g.shape[0] // 8
    t = S @ g + gt
    ta5 = t[:n]
    tb5 = t[n : 2 * n][::-1]
    tb6 = t[2 * n : 3 * n]
    tc6 = t[3 * n : 4 * n][::-1]
    tc7 = t[4 * n : 5 * n]
    td7 = t[5 * n : 6 * n][::-1]
    td8 = t[6 * n : 7 * n]
    ta8 = t[7 * n :][::-1]
    ga = np.concatenate([g[:n], ta5, ta8, g[7 * n :]])
    gb = np.concatenate([g[n : 3 * n], tb6, tb5])
    gc = np.concatenate([tc6, g[3 * n : 5 * n], tc7])
    gd = np.concatenate([td8, td7, g[5 * n : 7 * n]])
    return np.stack([ga, gb, gc, gd])


def down_pass(bdry, S_lst, g_lst, Y, v, kind="dtn"):
    """``down_pass_uniform_2D_{DtN,ItI}`` with multi-source data (nb, nsrc).

    Returns (u (n_leaves, n, nsrc), leaf boundary data (n_leaves, 4q, nsrc)).
    """
    fn = propagate_down_dtn if kind == "dtn" else propagate_down_iti
    g = bdry[None]
    for lvl in range(len(S_lst) - 1, -1, -1):
        S, gt = S_lst[lvl], g_lst[lvl]
        g = np.concatenate([fn(S[i], g[i], gt[i]) for i in range(S.shape[0])], axis=0)
    return Y @ g + v, g


# ---------------------------------------------------------------------------
# Up pass (no-source solver, new source)
# ---------------------------------------------------------------------------
def up_pass_iti(source, Phi, QH, D_inv_lst, BD_inv_lst, nb):
    """``up_pass_uniform_2D_ItI`` (``up_pass/_uniform_2D_ItI.py:8-219``):
    v = Phi f_int, h = QH v, then per level g_tilde = -D^{-1} h_int (rows
    permuted by r) and h_ext' = h_ext - B D^{-1} h_int (rolled).
    source: (n_leaves, n, nsrc).  Returns v, g_tilde_lst."""
    v = Phi @ source[:, nb:]
    h = QH @ v
    g_lst = []
    for Di, BDi in zip(D_inv_lst, BD_inv_lst):
        nm = Di.shape[0]
        new_h, gts = [], []
        for i in range(nm):
            ha, hb, hc, hd = h[4 * i], h[4 * i + 1], h[4 * i + 2], h[4 * i + 3]
            ne = ha.shape[0]
            m = ne // 4
            z = np.zeros((ne, ne), dtype=Di.dtype)
            _, _, _, _, _, _, h_int, h_ext = _iti_blocks(z, z, z, z, ha, hb, hc, hd)
            gt = -Di[i] @ h_int
            new_h.append(np.roll(h_ext - BDi[i] @ h_int, -m, axis=0))
            gts.append(gt[_iti_perm(m)])
        h = np.stack(new_h)
        g_lst.append(np.stack(gts))
    return v, g_lst
