"""Host-built index tables for the merge and down-pass kernels.

A merge level combines nc children whose exterior boundary data has nec
entries each.  The tables say, for every node of the merged problem, which
child and which local boundary index it comes from:

* ``int_rows`` (ni, 4): interface equation r sums the outgoing data of up to
  two (child, local) slots;
* ``int_cols`` (ni, 4): interface unknown j is the incoming data of up to two
  (child, local) slots;
* ``ext`` (ne, 4): merged exterior node k is (child, local) -- already in the
  merged node's final order;
* ``child_src`` (nc*nec,): for the down pass, where each child boundary slot
  takes its data from (>= 0: merged exterior index, < 0: interface -(j+1)).

2D layouts follow the reference exactly (sides S, E, N, W; children
a=SW, b=SE, c=NE, d=NW; interfaces 5=a|b, 6=b|c, 7=c|d, 8=d|a ordered from
the outer boundary inwards):
* DtN unknown/equation order [5, 6, 7, 8]
  (``merge/_uniform_2D_DtN.py:385-437`` index sets, ``:283-307`` blocks);
* ItI unknown order [a5, b5, b6, c6, c7, d7, d8, a8], the reference's order
  after its row permutation r (``merge/_uniform_2D_ItI.py:357-373``); the
  equation for unknown "x_k" is the outgoing data of its neighbour across
  the interface, so D = I + (off-diagonal blocks);
* exterior order = the reference's [a1, b2, c3, d4] rolled by -m
  (``merge/_uniform_2D_DtN.py:342-346``).
"""

from __future__ import annotations

import functools

import numpy as np
import torch

A_, B_, C_, D_ = 0, 1, 2, 3


def _iface_2D(m: int):
    """The four interfaces as lists of ((child, local), (child, local)) node
    pairs, each ordered from the outer boundary inwards."""
    j = np.arange(m)
    i5 = [((A_, m + k), (B_, 4 * m - 1 - k)) for k in j]
    i6 = [((B_, 2 * m + k), (C_, m - 1 - k)) for k in j]
    i7 = [((C_, 3 * m + k), (D_, 2 * m - 1 - k)) for k in j]
    i8 = [((D_, k), (A_, 3 * m - 1 - k)) for k in j]
    return i5, i6, i7, i8


def _ext_2D(m: int):
    """Merged exterior in final order: S (a_S, b_S), E (b_E, c_E), N (c_N, d_N), W (d_W, a_W)."""
    out = []
    out += [(A_, k) for k in range(0, m)]
    out += [(B_, k) for k in range(0, m)]
    out += [(B_, k) for k in range(m, 2 * m)]
    out += [(C_, k) for k in range(m, 2 * m)]
    out += [(C_, k) for k in range(2 * m, 3 * m)]
    out += [(D_, k) for k in range(2 * m, 3 * m)]
    out += [(D_, k) for k in range(3 * m, 4 * m)]
    out += [(A_, k) for k in range(3 * m, 4 * m)]
    return out


class MergeTables:
    """Index tables of one merge level.

    int_rows/int_cols are in the *internal* unknown order the factorisation
    uses; cols_final is int_cols in the final (stored) order and
    final_of_internal maps internal -> final rows (identity for DtN).  upd
    lists the nonzero (r0, nr, k0, nk) blocks of B (exterior x unknowns).
    """

    def __init__(self, kind, nc, nec, int_rows, int_cols, ext, child_src, cols_final=None, final_of_internal=None,
                 block=None):
        self.kind = kind
        self.nc = nc
        self.nec = nec
        self.ni = int_rows.shape[0]
        self.ne = ext.shape[0]
        self.int_rows = int_rows
        self.int_cols = int_cols
        self.ext = ext
        self.child_src = child_src
        self.cols_final = int_cols if cols_final is None else cols_final
        self.final_of_internal = (
            np.arange(self.ni, dtype=np.int32) if final_of_internal is None else final_of_internal.astype(np.int32)
        )
        self.upd = _sparse_blocks(self.ext, self.cols_final, block) if block else np.zeros((0, 4), np.int32)
        # ItI couplings: rectangles worth a product of their own once a block
        # spans at least one 64-row GEMM tile
        self.d12_blk = self.d21_blk = np.zeros((0, 4), np.int32)
        if kind == "iti" and block and block >= _COUPLING_MIN_BLOCK:
            nf = self.ni // 2
            self.d12_blk = _coupling_blocks(int_rows[:nf], int_cols[nf:], block)
            self.d21_blk = _coupling_blocks(int_rows[nf:], int_cols[:nf], block)
        self.adj_src = _adj_src(nc, nec, int_rows, ext)
        self._dev = {}

    def adj_device(self, dev):
        """adj_src on the device (the adjoint up pass's child-slot sources)."""
        key = ("adj", str(dev))
        if key not in self._dev:
            self._dev[key] = torch.from_numpy(np.ascontiguousarray(self.adj_src, dtype=np.int32)).to(dev)
        return self._dev[key]

    def device(self, dev):
        key = str(dev)
        if key not in self._dev:
            self._dev[key] = tuple(
                torch.from_numpy(np.ascontiguousarray(x, dtype=np.int32)).to(dev)
                for x in (self.int_rows, self.int_cols, self.ext, self.child_src, self.cols_final,
                          self.final_of_internal)
            )
        return self._dev[key]


def _children(entry):
    return {int(entry[0])} | ({int(entry[2])} if entry[2] >= 0 else set())


def _sparse_blocks(ext, cols_final, blk):
    """Nonzero blocks of B = gather(ext, cols_final) at block size blk, merged
    into rectangles: consecutive exterior blocks with the same set of unknown
    blocks share a rectangle; consecutive unknown blocks form one run."""
    ne, ni = ext.shape[0], cols_final.shape[0]
    assert ne % blk == 0 and ni % blk == 0
    eb = [set().union(*[_children(ext[i]) for i in range(b * blk, (b + 1) * blk)]) for b in range(ne // blk)]
    ub = [set().union(*[_children(cols_final[i]) for i in range(b * blk, (b + 1) * blk)]) for b in range(ni // blk)]
    sets = [tuple(u for u in range(len(ub)) if eb[e] & ub[u]) for e in range(len(eb))]
    rects = []
    e = 0
    while e < len(sets):
        e1 = e
        while e1 + 1 < len(sets) and sets[e1 + 1] == sets[e]:
            e1 += 1
        us = sets[e]
        i = 0
        while i < len(us):
            j = i
            while j + 1 < len(us) and us[j + 1] == us[j] + 1:
                j += 1
            rects.append((e * blk, (e1 - e + 1) * blk, us[i] * blk, (us[j] - us[i] + 1) * blk))
            i = j + 1
        e = e1 + 1
    return np.array(rects, dtype=np.int32).reshape(-1, 4)


_COUPLING_MIN_BLOCK = 64


def _coupling_blocks(rows, cols, blk):
    """Nonzero rectangles of the coupling block gather(rows, cols): entry
    (i, j) is nonzero only where an equation of rows[i] reads the child that
    unknown cols[j] feeds.  One rectangle per row block (its nonzero column
    blocks must form one run), consecutive row blocks with the same run merged;
    empty when a row block's columns do not form a single run."""
    nr, nk = rows.shape[0], cols.shape[0]
    if nr % blk or nk % blk:
        return np.zeros((0, 4), np.int32)
    rb = [set().union(*[_children(rows[i]) for i in range(b * blk, (b + 1) * blk)]) for b in range(nr // blk)]
    kb = [set().union(*[_children(cols[i]) for i in range(b * blk, (b + 1) * blk)]) for b in range(nk // blk)]
    runs = []
    for r in rb:
        ks = [k for k in range(len(kb)) if r & kb[k]]
        if not ks or ks != list(range(ks[0], ks[-1] + 1)):
            return np.zeros((0, 4), np.int32)
        runs.append((ks[0], ks[-1] + 1))
    rects = []
    b = 0
    while b < len(runs):
        e = b
        while e + 1 < len(runs) and runs[e + 1] == runs[b]:
            e += 1
        k0, k1 = runs[b]
        rects.append((b * blk, (e - b + 1) * blk, k0 * blk, (k1 - k0) * blk))
        b = e + 1
    return np.array(rects, dtype=np.int32).reshape(-1, 4)


def _pack(pairs):
    out = np.full((len(pairs), 4), -1, dtype=np.int32)
    out[:, 1] = 0
    out[:, 3] = 0
    for r, pr in enumerate(pairs):
        for a, (c, l) in enumerate(pr):
            out[r, 2 * a] = c
            out[r, 2 * a + 1] = l
    return out


def _adj_src(nc, nec, int_rows, ext):
    """For the adjoint of the up pass: each child slot feeds either one
    merged exterior node (index >= 0) or one interface equation row r of
    int_rows (internal order, coded -(r+1))."""
    src = np.full(nc * nec, np.iinfo(np.int32).min, dtype=np.int64)
    for e, entry in enumerate(ext):
        src[entry[0] * nec + entry[1]] = e
    for r, entry in enumerate(int_rows):
        for a in (0, 2):
            if entry[a] >= 0:
                slot = entry[a] * nec + entry[a + 1]
                assert src[slot] == np.iinfo(np.int32).min, "child slot feeds two equations"
                src[slot] = -(r + 1)
    assert (src != np.iinfo(np.int32).min).all(), "child slot feeds nothing"
    return src.astype(np.int32)


def _child_src(nc, nec, ext, unknown_slots):
    """unknown_slots[j] = list of (child, local) slots fed by unknown j."""
    src = np.full(nc * nec, np.iinfo(np.int32).min, dtype=np.int64)
    for k, (c, l) in enumerate(ext):
        src[c * nec + l] = k
    for j, slots in enumerate(unknown_slots):
        for (c, l) in slots:
            src[c * nec + l] = -(j + 1)
    assert (src != np.iinfo(np.int32).min).all(), "child slot without a source"
    return src.astype(np.int32)


@functools.lru_cache(maxsize=None)
def quad_merge_tables(m: int, kind: str) -> MergeTables:
    """Tables for a 2D 4:1 merge of children with m Gauss nodes per side."""
    nec = 4 * m
    i5, i6, i7, i8 = _iface_2D(m)
    ext = _ext_2D(m)
    if kind == "dtn":
        nodes = i5 + i6 + i7 + i8
        rows = _pack(nodes)
        cols = rows.copy()
        unknown_slots = [list(p) for p in nodes]
        ext_arr = _pack([[e] for e in ext])
        src = _child_src(4, nec, ext, unknown_slots)
        return MergeTables("dtn", 4, nec, rows, cols, ext_arr, src, block=m)
    elif kind == "iti":
        # (child, local) of the unknown and of its partner across the interface
        a5 = [(p[0], p[1]) for p in i5]
        b5 = [(p[1], p[0]) for p in i5]
        b6 = [(p[0], p[1]) for p in i6]
        c6 = [(p[1], p[0]) for p in i6]
        c7 = [(p[0], p[1]) for p in i7]
        d7 = [(p[1], p[0]) for p in i7]
        d8 = [(p[0], p[1]) for p in i8]
        a8 = [(p[1], p[0]) for p in i8]
        final = a5 + b5 + b6 + c6 + c7 + d7 + d8 + a8
        # internal order: unknowns of children a, c (F half) then b, d (G half);
        # the equation of every F unknown involves only G unknowns and vice
        # versa, so D = I + [[0, D12], [D21, 0]]
        internal = a5 + a8 + c6 + c7 + b5 + b6 + d7 + d8
        pos = {own: i for i, (own, _) in enumerate(final)}
        final_of_internal = np.array([pos[own] for own, _ in internal], dtype=np.int32)
        cols = _pack([[own] for own, _ in internal])
        rows = _pack([[partner] for _, partner in internal])
        cols_final = _pack([[own] for own, _ in final])
        ext_arr = _pack([[e] for e in ext])
        src = _child_src(4, nec, ext, [[own] for own, _ in final])
        return MergeTables("iti", 4, nec, rows, cols, ext_arr, src, cols_final=cols_final,
                           final_of_internal=final_of_internal, block=m)
    else:
        raise ValueError(kind)


# ---------------------------------------------------------------------------
# 3D 8:1 merges (merge/_uniform_3D_DtN.py:128-541, merge/_schur_complement.py:294-774)
# ---------------------------------------------------------------------------
# children a..h = 0..7 as in the oct subdivision (_grid_creation_3D.py:29-78):
# a..d upper z (x-y quadrants SW, SE, NE, NW), e..h lower z.
# faces 0..5 of a box: x-, x+, y-, y+, z-, z+ (the reference's faces 1..6).
# Interfaces 9..20 in the reference's unknown order: (child, face) on both sides.
_OCT_IFACES = (
    (0, 1, 1, 0), (1, 3, 2, 2), (2, 0, 3, 1), (3, 2, 0, 3),  # 9 a|b, 10 b|c, 11 c|d, 12 d|a
    (4, 1, 5, 0), (5, 3, 6, 2), (6, 0, 7, 1), (7, 2, 4, 3),  # 13 e|f, 14 f|g, 15 g|h, 16 h|e
    (0, 4, 4, 5), (1, 4, 5, 5), (2, 4, 6, 5), (3, 4, 7, 5),  # 17 a|e, 18 b|f, 19 c|g, 20 d|h
)
# merged face k = the children's face-k blocks in this order (the reference's
# get_rearrange_indices, :443-541): the quad order of the face's two coordinates
_OCT_EXT = ((4, 7, 3, 0), (5, 6, 2, 1), (4, 5, 1, 0), (7, 6, 2, 3), (4, 5, 6, 7), (0, 1, 2, 3))


@functools.lru_cache(maxsize=None)
def oct_merge_tables(n: int) -> MergeTables:
    """Tables for a 3D 8:1 DtN merge of children with n Gauss nodes per face.

    Both children see a shared face in the same (first, second coordinate)
    order, so interface node j is node j of each child's face block -- no
    flips in 3D."""
    nec = 6 * n
    nodes = [((c1, f1 * n + j), (c2, f2 * n + j)) for (c1, f1, c2, f2) in _OCT_IFACES for j in range(n)]
    rows = _pack(nodes)
    ext = [(c, k * n + j) for k, kids in enumerate(_OCT_EXT) for c in kids for j in range(n)]
    ext_arr = _pack([[e] for e in ext])
    src = _child_src(8, nec, ext, [list(p) for p in nodes])
    return MergeTables("dtn", 8, nec, rows, rows.copy(), ext_arr, src, block=n)
