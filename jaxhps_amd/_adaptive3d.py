"""Adaptive (non-uniform) 3D octrees: refinement, grids and the DtN build /
solve on them.

Follows ``_adaptive_discretization_3D.py``, ``_discretization_tree_operations_3D.py``,
``_grid_creation_3D.py:193-375`` (points), ``local_solve/_adaptive_3D_DtN.py``,
``merge/_adaptive_3D_DtN.py`` with ``merge/_utils_adaptive_3D_DtN.py`` and
``down_pass/_adaptive_3D_DtN.py``; the device design is the 2D one
(``_adaptive.py``): one batched leaf solve with per-leaf scaling, compression
T' = R T P of the finer side of each interface (R, P: block diagonals of
identities and the face-panel coarsening L_1f4 / refining L_4f1 operators on
each group of four panels), children packed into padded blocks, and the
ordinary index-table 8:1 merge, one batch per level and signature.

Faces 0..5 = x-, x+, y-, y+, z-, z+.  A node's face points run over its face
panels in the recursive quad order of the merged faces (children per face
as in ``_tables._OCT_EXT``), each panel's q^2 Gauss points with the first
face coordinate slowest.
"""

from __future__ import annotations


import numpy as np
import torch

from ._adaptive import _pack
from ._discretization_tree import DiscretizationNode3D, get_all_leaves
from ._grid_creation import bounds_to_cheby_points_3D, bounds_to_gauss_face, rearrange_indices_ext_int_3D
from ._tables import _OCT_EXT, _OCT_IFACES
from .quadrature import (
    barycentric_lagrange_interpolation_matrix_1D,
    chebyshev_points,
    chebyshev_weights,
    gauss_points,
)

# (fixed axis, take the upper bound?, first face coordinate, second) per face
_FACE_AXES = ((0, False, 1, 2), (0, True, 1, 2), (1, False, 0, 2), (1, True, 0, 2), (2, False, 0, 1), (2, True, 0, 1))


# ---------------------------------------------------------------------------
# tree operations (_discretization_tree_operations_3D.py)
# ---------------------------------------------------------------------------
def _b(nd):
    return (nd.xmin, nd.xmax, nd.ymin, nd.ymax, nd.zmin, nd.zmax)


def get_eight_children(parent):
    """Children a..h: the four x-y quadrants (SW, SE, NE, NW) of the upper z
    half, then of the lower z half."""
    x0, x1, y0, y1, z0, z1 = _b(parent)
    xm, ym, zm = (x0 + x1) / 2, (y0 + y1) / 2, (z0 + z1) / 2
    quads = ((x0, xm, y0, ym), (xm, x1, y0, ym), (xm, x1, ym, y1), (x0, xm, ym, y1))
    d = parent.depth + 1
    out = [DiscretizationNode3D(*qd, zm, z1, depth=d) for qd in quads]
    out += [DiscretizationNode3D(*qd, z0, zm, depth=d) for qd in quads]
    return tuple(out)


def _child_index(cur, node):
    c = [(node.xmin + node.xmax) / 2, (node.ymin + node.ymax) / 2, (node.zmin + node.zmax) / 2]
    xm, ym, zm = (cur.xmin + cur.xmax) / 2, (cur.ymin + cur.ymax) / 2, (cur.zmin + cur.zmax) / 2
    quad = (0 if c[1] < ym else 3) if c[0] < xm else (1 if c[1] < ym else 2)
    return quad + (4 if c[2] < zm else 0)


def _face_counts(nd, q):
    """Gauss points on faces 0..5 = q^2 x the leaves on each face."""
    return tuple(q * q * len(_face_leaves(nd, f)) for f in range(6))


def _set_counts(nd, q):
    for f, v in enumerate(_face_counts(nd, q)):
        setattr(nd, f"n_{f}", v)


def add_eight_children(add_to, root=None, q: int | None = None) -> None:
    """Split a leaf into eight children; with root and q keep the face counts
    n_0..n_5 of add_to and its ancestors current
    (``_discretization_tree_operations_3D.py:26-99``)."""
    if len(add_to.children):
        return
    add_to.children = get_eight_children(add_to)
    if q is None:
        return
    for ch in add_to.children:
        _set_counts(ch, q)
    _set_counts(add_to, q)
    if root is None or root is add_to:
        return
    cur = root
    path = []
    while cur is not add_to:
        if not cur.children:
            raise ValueError("node is not in the tree under root")
        path.append(cur)
        cur = cur.children[_child_index(cur, add_to)]
    for nd in path:
        _set_counts(nd, q)


def _face_leaves(nd, f):
    """Leaves on face f of nd in the face's panel order."""
    if not len(nd.children):
        return [nd]
    return [lf for c in _OCT_EXT[f] for lf in _face_leaves(nd.children[c], f)]


def get_ordered_lst_of_boundary_nodes(root):
    """Leaves on the root's faces 1..6 in panel order (``_grid_creation_3D.py:375-417``)."""
    return tuple(_face_leaves(root, f) for f in range(6))


def assign_face_counts(root, q: int) -> None:
    def rec(nd):
        for ch in nd.children:
            rec(ch)
        _set_counts(nd, q)

    rec(root)


# ---------------------------------------------------------------------------
# grids
# ---------------------------------------------------------------------------
def compute_interior_Chebyshev_points_adaptive_3D(root, p: int) -> np.ndarray:
    c = chebyshev_points(p)
    return np.stack([bounds_to_cheby_points_3D(np.array(_b(lf)), c) for lf in get_all_leaves(root)])


def compute_boundary_Gauss_points_adaptive_3D(root, q: int) -> np.ndarray:
    g = gauss_points(q)
    rb = _b(root)
    out = []
    for f, (ax, upper, a1, a2) in enumerate(_FACE_AXES):
        for lf in _face_leaves(root, f):
            lb = _b(lf)
            fp = bounds_to_gauss_face(np.array([lb[2 * a1], lb[2 * a1 + 1], lb[2 * a2], lb[2 * a2 + 1]]), g)
            xyz = np.zeros((fp.shape[0], 3))
            xyz[:, ax] = rb[2 * ax + 1] if upper else rb[2 * ax]
            xyz[:, a1] = fp[:, 0]
            xyz[:, a2] = fp[:, 1]
            out.append(xyz)
    return np.concatenate(out)


# ---------------------------------------------------------------------------
# refinement (_adaptive_discretization_3D.py)
# ---------------------------------------------------------------------------
def precompute_L_8f1(p: int) -> np.ndarray:
    """(8 p^3, p^3) interpolation from a leaf's grid to its eight children's
    grids (a..h), every grid in leaf point order."""
    c = chebyshev_points(p)
    r = rearrange_indices_ext_int_3D(p)
    ix, rem = np.divmod(r, p * p)
    iy, iz = np.divmod(rem, p)
    X, Y, Z = c[ix], c[iy], c[iz]
    blocks = []
    for cb in get_eight_children(DiscretizationNode3D(-1.0, 1.0, -1.0, 1.0, -1.0, 1.0)):
        cx, cy, cz = (cb.xmin + cb.xmax) / 2, (cb.ymin + cb.ymax) / 2, (cb.zmin + cb.zmax) / 2
        Lx = barycentric_lagrange_interpolation_matrix_1D(c, cx + 0.5 * X)
        Ly = barycentric_lagrange_interpolation_matrix_1D(c, cy + 0.5 * Y)
        Lz = barycentric_lagrange_interpolation_matrix_1D(c, cz + 0.5 * Z)
        blocks.append(Lx[:, ix] * Ly[:, iy] * Lz[:, iz])
    return np.concatenate(blocks, axis=0)


def get_squared_l2_norm_single_voxel(f_evals, bounds, p: int) -> float:
    """Clenshaw-Curtis estimate of the squared L2 norm of f on one voxel
    (``_adaptive_discretization_3D.py:441-473``)."""
    b = np.asarray(bounds, dtype=np.float64)
    w = np.outer(np.outer(chebyshev_weights(p, b[:2]), chebyshev_weights(p, b[2:4])).ravel(),
                 chebyshev_weights(p, b[4:])).ravel()
    return float(np.sum(np.asarray(f_evals) ** 2 * w[rearrange_indices_ext_int_3D(p)]))


def _voxel_points(B, p):
    c = chebyshev_points(p)
    return np.stack([bounds_to_cheby_points_3D(b, c) for b in B])


def _kids(B):
    return np.array([[_b(k) for k in get_eight_children(DiscretizationNode3D(*b))] for b in B])


def volumes_to_check(newly_refined, root):
    """Same-size boxes sharing a face with newly_refined, inside the root
    (``_adaptive_discretization_3D.py:213-314``)."""
    s = newly_refined.xmax - newly_refined.xmin
    b = list(_b(newly_refined))
    rb = _b(root)
    out = []
    for ax in range(3):
        lo, hi = b[2 * ax], b[2 * ax + 1]
        if lo - s >= rb[2 * ax]:
            v = b.copy()
            v[2 * ax], v[2 * ax + 1] = lo - s, lo
            out.append(tuple(v))
        if hi + s <= rb[2 * ax + 1]:
            v = b.copy()
            v[2 * ax], v[2 * ax + 1] = hi, hi + s
            out.append(tuple(v))
    return out


def find_or_add_child(node, root, q, xmin, xmax, ymin, ymax, zmin, zmax):
    """Make the octree cell exist below node by splitting the leaf covering it;
    returns the split node or None (``_adaptive_discretization_3D.py:317-421``)."""
    want = xmax - xmin
    side = node.xmax - node.xmin
    if np.isclose(2 * want, side, rtol=1e-5, atol=1e-8):
        if not len(node.children):
            add_eight_children(node, root=root, q=q)
            return node
        return None
    box = DiscretizationNode3D(xmin, xmax, ymin, ymax, zmin, zmax)
    child = node.children[_child_index(node, box)]
    if np.isclose(side, 4 * want, rtol=1e-5, atol=1e-8):
        if not len(child.children):
            add_eight_children(child, root=root, q=q)
            return child
        return None
    if side < 2 * want:
        raise ValueError("Requested volume is too large for the current node")
    return find_or_add_child(child, root, q, xmin, xmax, ymin, ymax, zmin, zmax)


def generate_adaptive_mesh_level_restriction(root, f_fn, tol: float, p: int, q: int, restrict_bool: bool = True,
                                             l2_norm: bool = False) -> None:
    """The 3D refinement loop of the 2D one (``_adaptive.generate_adaptive_mesh_level_restriction_2D``)
    with eight children and face neighbours."""
    R = precompute_L_8f1(p)
    m = p**3
    if l2_norm:
        add_eight_children(root, root=root, q=q)
        lv = get_all_leaves(root)
        fv = np.asarray(f_fn(_voxel_points([np.array(_b(x)) for x in lv], p)))
        gnorm = sum(get_squared_l2_norm_single_voxel(fv[i], _b(x), p) for i, x in enumerate(lv))
        tol = tol**2
    else:
        gnorm = float(np.max(np.asarray(f_fn(_voxel_points(_kids([_b(root)])[0], p)))))
    queue = list(get_all_leaves(root))
    while queue:
        B = np.array([_b(x) for x in queue])
        K = _kids(B)
        f0 = np.asarray(f_fn(_voxel_points(B, p)))
        f1 = np.asarray(f_fn(_voxel_points(K.reshape(-1, 6), p))).reshape(len(queue), 8 * m)
        diff = f0 @ R.T - f1
        if l2_norm:
            err = np.array([sum(get_squared_l2_norm_single_voxel(diff[i, k * m:(k + 1) * m], K[i, k], p)
                                for k in range(8)) for i in range(len(queue))])
            ok = err / gnorm < tol
        else:
            c = np.maximum(gnorm, np.max(np.abs(f1), axis=1))
            ok = np.max(np.abs(diff), axis=1) / c < tol
            gnorm = float(np.max(c))
        nxt = []
        for nd, good in zip(queue, ok):
            if good:
                continue
            add_eight_children(nd, root=root, q=q)
            nxt.extend(nd.children)
            if restrict_bool:
                check_q = [nd]
                while check_q:
                    cur = check_q.pop()
                    for box in volumes_to_check(cur, root):
                        made = find_or_add_child(root, root, q, *box)
                        if made is not None:
                            check_q.append(made)
                            nxt.extend(made.children)
        queue = nxt


# ---------------------------------------------------------------------------
# merge plans
# ---------------------------------------------------------------------------
def precompute_projection_ops_3D(q: int):
    """L_4f1 (4q^2 x q^2): a face panel's q x q Gauss grid to its quadrants
    A, B, C, D = (lo, lo), (hi, lo), (hi, hi), (lo, hi); L_1f4 (q^2 x 4q^2):
    each coarse point from the quadrant holding it (q even)
    (``_precompute_operators_3D.py:323-409``)."""
    if q % 2:
        raise ValueError("q must be even.")
    g = gauss_points(q)
    lo, hi = 0.5 * g - 0.5, 0.5 * g + 0.5
    I = barycentric_lagrange_interpolation_matrix_1D
    quads = ((lo, lo, True, True), (hi, lo, False, True), (hi, hi, False, False), (lo, hi, True, False))
    L41 = np.concatenate([np.kron(I(g, a), I(g, b)) for a, b, _, _ in quads])
    L14 = np.zeros((q * q, 4 * q * q))
    half = np.arange(q) < q // 2
    for k, (a, b, alo, blo) in enumerate(quads):
        ia = half if alo else ~half
        ib = half if blo else ~half
        rows = np.nonzero(np.outer(ia, ib).ravel())[0]
        L14[rows, k * q * q:(k + 1) * q * q] = np.kron(I(a, g[ia]), I(b, g[ib]))
    return L41, L14


def _walk(l0, l1):
    o0, o1 = [], []
    i = j = 0
    while i < len(l0) and j < len(l1):
        s0, s1 = l0[i].xmax - l0[i].xmin, l1[j].xmax - l1[j].xmin
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


def find_projection_lists_3D(node):
    """Per interface 9..20: panel compression flags of its two children's
    faces (``merge/_utils_adaptive_3D_DtN.py:11-166``)."""
    ch = node.children
    return [_walk(_face_leaves(ch[c1], f1), _face_leaves(ch[c2], f2)) for (c1, f1, c2, f2) in _OCT_IFACES]


class _Plan3:
    """Tables and compression operators of the 8:1 merges of one signature."""

    kind = "dtn"
    nc = 8
    upd = np.zeros((0, 4), np.int32)
    d12_blk = np.zeros((0, 4), np.int32)
    d21_blk = np.zeros((0, 4), np.int32)

    def __init__(self, counts, flags, q, L41, L14):
        q2 = q * q
        fl = [[[False] * (counts[c][f] // q2) for f in range(6)] for c in range(8)]
        for i, (c1, f1, c2, f2) in enumerate(_OCT_IFACES):
            fl[c1][f1], fl[c2][f2] = list(flags[i][0]), list(flags[i][1])
        self.n = [sum(c) for c in counts]
        self.R, self.P, off, length = [], [], [], []
        for c in range(8):
            rows, col, blocks, so, sl = 0, 0, [], [], []
            for f in range(6):
                so.append(rows)
                fls = fl[c][f]
                i = 0
                while i < len(fls):
                    if fls[i]:
                        blocks.append((rows, col, True))
                        col += 4 * q2
                        i += 4
                    else:
                        blocks.append((rows, col, False))
                        col += q2
                        i += 1
                    rows += q2
                sl.append(rows - so[-1])
            assert col == self.n[c]
            off.append(so)
            length.append(sl)
            if rows == self.n[c]:
                self.R.append(None)
                self.P.append(None)
                continue
            R = np.zeros((rows, self.n[c]))
            P = np.zeros((self.n[c], rows))
            for r0, c0, comp in blocks:
                if comp:
                    R[r0:r0 + q2, c0:c0 + 4 * q2] = L14
                    P[c0:c0 + 4 * q2, r0:r0 + q2] = L41
                else:
                    R[r0:r0 + q2, c0:c0 + q2] = np.eye(q2)
                    P[c0:c0 + q2, r0:r0 + q2] = np.eye(q2)
            self.R.append(R)
            self.P.append(P)
        self.m = [sum(length[c]) for c in range(8)]
        self.nmax = max(self.m)
        pairs = []
        for (c1, f1, c2, f2) in _OCT_IFACES:
            L = length[c1][f1]
            assert L == length[c2][f2], "interface discretisations do not match"
            pairs += [((c1, off[c1][f1] + t), (c2, off[c2][f2] + t)) for t in range(L)]
        ext = [(c, off[c][k] + t) for k, kids in enumerate(_OCT_EXT) for c in kids for t in range(length[c][k])]
        self.ni, self.ne = len(pairs), len(ext)
        self.int_rows = _pack(pairs)
        self.int_cols = self.int_rows.copy()
        self.ext = _pack([[e] for e in ext])
        src = np.zeros(8 * self.nmax, dtype=np.int32)
        for k, (c, l) in enumerate(ext):
            src[c * self.nmax + l] = k
        for j, pr in enumerate(pairs):
            for c, l in pr:
                src[c * self.nmax + l] = -(j + 1)
        self.child_src = src
        self._dev = {}

    @property
    def nec(self):
        return self.nmax

    @property
    def cols_final(self):
        return self.int_cols

    @property
    def final_of_internal(self):
        return np.arange(self.ni, dtype=np.int32)

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


def _signature3(node, q):
    counts = tuple(_face_counts(ch, q) for ch in node.children)
    flags = tuple((tuple(a), tuple(b)) for a, b in find_projection_lists_3D(node))
    return counts, flags


# ---------------------------------------------------------------------------
# stages
# ---------------------------------------------------------------------------
def local_solve_stage_adaptive_3D_DtN(pde_problem, device=None, host_device=None):
    """Leaf DtN maps for leaves of different sizes (``local_solve/_adaptive_3D_DtN.py:13-128``)."""
    from ._adaptive import local_solve_stage_adaptive_2D_DtN

    return local_solve_stage_adaptive_2D_DtN(pde_problem, device=device, host_device=host_device)


def merge_stage_adaptive_3D_DtN(pde_problem, T_leaf=None, h_leaf=None) -> None:
    """8:1 merges from the deepest level to the root (``merge/_adaptive_3D_DtN.py:30-149``)."""
    from ._adaptive import adaptive_merge_stage

    L41, L14 = pde_problem.L_4f1, pde_problem.L_1f4
    q = pde_problem.domain.q
    adaptive_merge_stage(pde_problem, T_leaf, h_leaf, _signature3, lambda sig: _Plan3(sig[0], sig[1], q, L41, L14))


def down_pass_adaptive_3D_DtN(pde_problem, boundary_data):
    """Root Dirichlet data (the six faces' samples as a list, or one array)
    down to the leaves (``down_pass/_adaptive_3D_DtN.py:15-142``)."""
    from ._adaptive import down_pass_adaptive_2D_DtN

    return down_pass_adaptive_2D_DtN(pde_problem, boundary_data)
