"""ORACLE (test infrastructure only) -- spectral setup built independently.

Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` may import this package; it is the checker, never the product.

The leaf operators here are constructed from point coordinates and textbook
formulas rather than from the index arithmetic the reference uses, so that a
mistake in either construction shows up as a mismatch:

* D-matrix from the barycentric derivative formula D_ij = (w_j/w_i)/(x_i-x_j)
  (Berrut & Trefethen, SIAM Review 2004, eq. 9.4) instead of the c-vector
  recipe of ``quadrature/_differentiation.py:9-47``.
* Lagrange interpolation by explicit products instead of the barycentric
  formula of ``quadrature/_interpolation.py:18-114``.
* The boundary-first leaf ordering as a counter-clockwise walk from the SW
  corner, derived from coordinates (``_grid_creation_2D.py:204-233`` encodes
  the same walk as index arithmetic).
* P/Q/G/QH from "which side does this point lie on" geometry
  (``_precompute_operators_2D.py:48-265``).
"""

from __future__ import annotations

import numpy as np


def cheb_points(n: int) -> np.ndarray:
    """cos(pi*(n-1-i)/(n-1)), ascending (``quadrature/_discretization.py:13-37``)."""
    return -np.cos(np.pi * np.arange(n) / (n - 1))


def gauss_nodes(n: int) -> np.ndarray:
    """Gauss-Legendre nodes via numpy (``quadrature/_discretization.py:105-116``)."""
    return np.polynomial.legendre.leggauss(n)[0]


def bary_weights(x: np.ndarray) -> np.ndarray:
    n = len(x)
    w = np.ones(n)
    for j in range(n):
        for k in range(n):
            if k != j:
                w[j] /= x[j] - x[k]
    return w


def diff_matrix(x: np.ndarray) -> np.ndarray:
    """Spectral differentiation matrix on nodes x (barycentric derivative form)."""
    n = len(x)
    w = bary_weights(x)
    D = np.zeros((n, n))
    for i in range(n):
        for j in range(n):
            if i != j:
                D[i, j] = (w[j] / w[i]) / (x[i] - x[j])
        D[i, i] = -np.sum(D[i, :])
    return D


def lagrange_matrix(from_x: np.ndarray, to_x: np.ndarray) -> np.ndarray:
    """M[i, j] = l_j(to_x[i]), l_j the Lagrange basis on from_x (explicit products)."""
    n = len(from_x)
    M = np.ones((len(to_x), n))
    for j in range(n):
        for k in range(n):
            if k != j:
                M[:, j] *= (to_x - from_x[k]) / (from_x[j] - from_x[k])
    return M


# ---------------------------------------------------------------------------
# 2D leaf ordering from coordinates
# ---------------------------------------------------------------------------
def leaf_order_2D(p: int):
    """Return (perm, X, Y) for the reference 2D leaf ordering.

    Grid: meshgrid(x, flipud(y), indexing="ij") flattened
    (``_grid_creation_2D.py:53-68``); perm lists the boundary as a CCW walk
    starting at the SW corner, then the interior in ascending flat index.
    """
    x = cheb_points(p)
    X, Y = np.meshgrid(x, x[::-1], indexing="ij")
    X = X.reshape(-1)
    Y = Y.reshape(-1)
    lo, hi = x[0], x[-1]
    flat = np.arange(p * p)

    def at(xv, yv):
        k = np.nonzero((X == xv) & (Y == yv))[0]
        assert len(k) == 1
        return int(k[0])

    walk = []
    # S side: y = lo, x increasing (includes both corners)
    walk += [at(xv, lo) for xv in x]
    # E side: x = hi, y increasing, skip SE
    walk += [at(hi, yv) for yv in x[1:]]
    # N side: y = hi, x decreasing, skip NE, include NW
    walk += [at(xv, hi) for xv in x[::-1][1:]]
    # W side: x = lo, y decreasing, skip NW and SW
    walk += [at(lo, yv) for yv in x[::-1][1:-1]]
    in_walk = set(walk)
    interior = [k for k in flat if k not in in_walk]
    perm = np.array(walk + interior)
    return perm, X[perm], Y[perm]


def leaf_ops_2D(p: int, q: int, half_side: float, eta: float | None = None):
    """All single-leaf operators on the reference's [-1,1]^2-scaled leaf.

    Returns a dict with D_x, D_y, D_xx, D_yy, D_xy (p^2 x p^2), and
    P (4(p-1) x 4q), Q (4q x p^2) for DtN; P_iti, G, QH for ItI when eta given.
    """
    perm, X, Y = leaf_order_2D(p)
    x = cheb_points(p)
    D1 = diff_matrix(x) / half_side
    n = p * p
    # Build d/dx and d/dy on the natural tensor grid from coordinates, then
    # express in the permuted order.
    # natural flat index: i*p + j with x_i, y = x[::-1][j]
    Dx_nat = np.kron(D1, np.eye(p))
    # y is reversed along j, so d/dy = -d/d(j-direction in ascending x)
    Dy_nat = -np.kron(np.eye(p), D1)
    Dx = Dx_nat[perm][:, perm]
    Dy = Dy_nat[perm][:, perm]
    ops = dict(D_x=Dx, D_y=Dy, D_xx=Dx @ Dx, D_yy=Dy @ Dy, D_xy=Dx @ Dy)

    g = gauss_nodes(q)
    lo, hi = x[0], x[-1]
    # Side point lists (indices into the permuted leaf order) in side order,
    # corners included on both adjacent sides.
    pos = {}
    for k in range(n):
        pos[(X[k], Y[k])] = k
    sides = [
        [pos[(xv, lo)] for xv in x],  # S, W->E
        [pos[(hi, yv)] for yv in x],  # E, S->N
        [pos[(xv, hi)] for xv in x[::-1]],  # N, E->W
        [pos[(lo, yv)] for yv in x[::-1]],  # W, N->S
    ]
    normals = [(0.0, -1.0), (1.0, 0.0), (0.0, 1.0), (-1.0, 0.0)]
    # Gauss nodes along each side follow the same direction as the side list.
    # Cheb side coordinate runs along x ascending for the list construction;
    # for reversed sides the parameter is reversed too, so the same 1D maps
    # apply in the side's own direction.
    P1 = lagrange_matrix(g, x)  # Gauss -> Cheb (p x q)
    Q1 = lagrange_matrix(x, g)  # Cheb -> Gauss (q x p)
    nb = 4 * (p - 1)

    # DtN P: every boundary Cheb point interpolated from each side it lies on,
    # averaged over sides (corners: 2 sides).
    P = np.zeros((nb, 4 * q))
    cnt = np.zeros(nb)
    for s in range(4):
        for a, k in enumerate(sides[s]):
            P[k, s * q : (s + 1) * q] += P1[a]
            cnt[k] += 1
    P /= cnt[:, None]

    # Q: normal derivative along each side (Cheb points incl. corners),
    # then interpolated to the side's Gauss points.
    Q = np.zeros((4 * q, n))
    Nrows = []
    for s in range(4):
        nx, ny = normals[s]
        rows = np.array([nx * Dx[k] + ny * Dy[k] for k in sides[s]])
        Nrows.append(rows)
        Q[s * q : (s + 1) * q] = Q1 @ rows
    ops["P"] = P
    ops["Q"] = Q

    if eta is not None:
        # ItI: each Cheb boundary point belongs to exactly one side, the side
        # whose walk segment contains it (corner -> the side that starts there).
        walk_sides = []
        for s in range(4):
            walk_sides.append(sides[s][:-1])  # drop the end corner
        P_iti = np.zeros((nb, 4 * q))
        G = np.zeros((nb, n), dtype=np.complex128)
        for s in range(4):
            nx, ny = normals[s]
            for a, k in enumerate(walk_sides[s]):
                P_iti[k, s * q : (s + 1) * q] = P1[a]
                G[k] = nx * Dx[k] + ny * Dy[k]
                G[k, k] += 1j * eta
        QH = np.zeros((4 * q, n), dtype=np.complex128)
        for s in range(4):
            rows = Nrows[s].astype(np.complex128)
            for a, k in enumerate(sides[s]):
                rows[a, k] -= 1j * eta
            QH[s * q : (s + 1) * q] = Q1 @ rows
        ops["P_iti"] = P_iti
        ops["G"] = G
        ops["QH"] = QH
    return ops


def quad_children(b):
    xmin, xmax, ymin, ymax = b
    xm, ym = (xmin + xmax) / 2, (ymin + ymax) / 2
    return [
        (xmin, xm, ymin, ym),
        (xm, xmax, ymin, ym),
        (xm, xmax, ym, ymax),
        (xmin, xm, ym, ymax),
    ]


def uniform_leaves_2D(bounds, L):
    boxes = [tuple(bounds)]
    for _ in range(L):
        boxes = [c for b in boxes for c in quad_children(b)]
    return boxes


def leaf_points_2D(bounds, L, p):
    perm, X, Y = leaf_order_2D(p)
    out = []
    for b in uniform_leaves_2D(bounds, L):
        cx, cy = (b[0] + b[1]) / 2, (b[2] + b[3]) / 2
        hx, hy = (b[1] - b[0]) / 2, (b[3] - b[2]) / 2
        out.append(np.stack([cx + hx * X, cy + hy * Y], -1))
    return np.stack(out)


def boundary_gauss_points_2D(bounds, L, q):
    """Root boundary Gauss points S(W->E), E(S->N), N(E->W), W(N->S)."""
    xmin, xmax, ymin, ymax = bounds
    g = gauss_nodes(q)
    n = 2**L
    def panels(a, b):
        br = np.linspace(a, b, n + 1)
        return np.concatenate([(br[i] + br[i + 1]) / 2 + (br[i + 1] - br[i]) / 2 * g for i in range(n)])
    xs = panels(xmin, xmax)
    ys = panels(ymin, ymax)
    return np.concatenate(
        [
            np.stack([xs, np.full_like(xs, ymin)], -1),
            np.stack([np.full_like(ys, xmax), ys], -1),
            np.stack([xs[::-1], np.full_like(xs, ymax)], -1),
            np.stack([np.full_like(ys, xmin), ys[::-1]], -1),
        ]
    )


def leaf_gauss_points_2D(b, q):
    """A leaf's 4q Gauss boundary points in its side order."""
    g = gauss_nodes(q)
    xmin, xmax, ymin, ymax = b
    xs = (xmin + xmax) / 2 + (xmax - xmin) / 2 * g
    ys = (ymin + ymax) / 2 + (ymax - ymin) / 2 * g
    return np.concatenate(
        [
            np.stack([xs, np.full_like(xs, ymin)], -1),
            np.stack([np.full_like(ys, xmax), ys], -1),
            np.stack([xs[::-1], np.full_like(xs, ymax)], -1),
            np.stack([np.full_like(ys, xmin), ys[::-1]], -1),
        ]
    )
