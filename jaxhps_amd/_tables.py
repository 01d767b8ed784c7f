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
    def __init__(self, nc, nec, int_rows, int_cols, ext, child_src):
        self.nc = nc
        self.nec = nec
        self.ni = int_rows.shape[0]
        self.ne = ext.shape[0]
        self.int_rows = int_rows
        self.int_cols = int_cols
        self.ext = ext
        self.child_src = child_src
        self._dev = {}

    def device(self, dev):
        key = str(dev)
        if key not in self._dev:
            self._dev[key] = tuple(
                torch.from_numpy(np.ascontiguousarray(x, dtype=np.int32)).to(dev)
                for x in (self.int_rows, self.int_cols, self.ext, self.child_src)
            )
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
        order = a5 + b5 + b6 + c6 + c7 + d7 + d8 + a8
        cols = _pack([[own] for own, _ in order])
        rows = _pack([[partner] for _, partner in order])
        unknown_slots = [[own] for own, _ in order]
    else:
        raise ValueError(kind)
    ext_arr = _pack([[e] for e in ext])
    src = _child_src(4, nec, ext, unknown_slots)
    return MergeTables(4, nec, rows, cols, ext_arr, src)
