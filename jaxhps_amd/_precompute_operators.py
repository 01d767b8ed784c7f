"""Leaf differentiation / interpolation operators (host, float64 numpy).

Every uniform leaf is the same square/cube, so one set of operators serves
all leaves.  Conventions follow ``_precompute_operators_2D.py:16-265`` and
``_precompute_operators_3D.py:17-320`` of the reference.
"""

from __future__ import annotations

import numpy as np

from ._grid_creation import rearrange_indices_ext_int_2D, rearrange_indices_ext_int_3D
from .quadrature import (
    barycentric_lagrange_interpolation_matrix_1D,
    barycentric_lagrange_interpolation_matrix_2D,
    chebyshev_points,
    differentiation_matrix_1D,
    gauss_points,
)


# --------------------------------------------------------------------------
# 2D
# --------------------------------------------------------------------------
def precompute_diff_operators_2D(p: int, half_side_len: float):
    """(D_x, D_y, D_xx, D_yy, D_xy) on the permuted leaf grid.

    D_y carries a minus sign because the y axis of the leaf grid runs
    north-to-south (``_precompute_operators_2D.py:16-44``); D_xy = D_x D_y.
    """
    r = rearrange_indices_ext_int_2D(p)
    D = differentiation_matrix_1D(chebyshev_points(p)) / half_side_len
    I = np.eye(p)
    dx = np.kron(D, I)[r][:, r]
    dy = -np.kron(I, D)[r][:, r]
    return dx, dy, dx @ dx, dy @ dy, dx @ dy


def precompute_P_2D_DtN(p: int, q: int) -> np.ndarray:
    """Gauss (4q) -> Chebyshev boundary (4(p-1)) interpolation with the four
    corner rows averaged between their two sides (:48-79)."""
    P = barycentric_lagrange_interpolation_matrix_1D(gauss_points(q), chebyshev_points(p))
    nb = 4 * (p - 1)
    out = np.zeros((nb, 4 * q))
    out[0:p, 0:q] = P
    out[p - 1 : 2 * p - 1, q : 2 * q] = P
    out[2 * (p - 1) : 3 * p - 2, 2 * q : 3 * q] = P
    out[3 * (p - 1) :, 3 * q :] = P[:-1]
    out[0, 3 * q :] = P[-1]
    for row in (0, p - 1, 2 * (p - 1), 3 * (p - 1)):
        out[row] /= 2.0
    return out


def precompute_P_2D_ItI(p: int, q: int) -> np.ndarray:
    """kron(I_4, P0) with P0 the 1D Gauss->Cheb map minus its last row (:83-99)."""
    P = barycentric_lagrange_interpolation_matrix_1D(gauss_points(q), chebyshev_points(p))
    return np.kron(np.eye(4), P[:-1])


def precompute_N_matrix_2D(du_dx: np.ndarray, du_dy: np.ndarray, p: int) -> np.ndarray:
    """Outward normal derivative on the 4p boundary points, corners counted
    once per side (:132-163)."""
    N = np.zeros((4 * p, p * p))
    N[:p] = -du_dy[:p]
    N[p : 2 * p] = du_dx[p - 1 : 2 * p - 1]
    N[2 * p : 3 * p] = du_dy[2 * p - 2 : 3 * p - 2]
    N[3 * p :] = -du_dx[3 * p - 3 : 4 * p - 3]
    N[-1] = -du_dx[0]
    return N


def precompute_N_tilde_matrix_2D(du_dx: np.ndarray, du_dy: np.ndarray, p: int) -> np.ndarray:
    """Normal derivative on the 4(p-1) boundary points, corners once (:166-192)."""
    m = p - 1
    N = np.zeros((4 * m, p * p))
    N[:m] = -du_dy[:m]
    N[m : 2 * m] = du_dx[m : 2 * m]
    N[2 * m : 3 * m] = du_dy[2 * m : 3 * m]
    N[3 * m :] = -du_dx[3 * m : 4 * m]
    return N


def precompute_Q_2D_DtN(p: int, q: int, du_dx: np.ndarray, du_dy: np.ndarray) -> np.ndarray:
    """Leaf solution -> outward normal derivative on the 4q Gauss points (:102-129)."""
    N = precompute_N_matrix_2D(du_dx, du_dy, p)
    Q1 = barycentric_lagrange_interpolation_matrix_1D(chebyshev_points(p), gauss_points(q))
    return np.kron(np.eye(4), Q1) @ N


def precompute_QH_2D_ItI(N: np.ndarray, p: int, q: int, eta: float) -> np.ndarray:
    """Leaf solution -> outgoing impedance u_n - i eta u on the Gauss points (:195-240)."""
    H = N.astype(np.complex128)
    ie = 1j * eta
    H[:p, :p] -= ie * np.eye(p)
    H[p : 2 * p, p - 1 : 2 * p - 1] -= ie * np.eye(p)
    H[2 * p : 3 * p, 2 * p - 2 : 3 * p - 2] -= ie * np.eye(p)
    H[3 * p : 4 * p - 1, 3 * p - 3 : 4 * p - 4] -= ie * np.eye(p - 1)
    H[4 * p - 1, 0] -= ie
    Q1 = barycentric_lagrange_interpolation_matrix_1D(chebyshev_points(p), gauss_points(q))
    return np.kron(np.eye(4), Q1) @ H


def precompute_G_2D_ItI(N_tilde: np.ndarray, eta: float) -> np.ndarray:
    """Leaf solution -> incoming impedance u_n + i eta u on the 4(p-1) Chebyshev
    boundary points (:243-265)."""
    G = N_tilde.astype(np.complex128)
    nb = G.shape[0]
    G[:nb, :nb] += 1j * eta * np.eye(nb)
    return G


# --------------------------------------------------------------------------
# 3D
# --------------------------------------------------------------------------
def precompute_diff_operators_3D(p: int, half_side_len: float):
    """(D_x, D_y, D_z, D_xx, D_yy, D_zz, D_xy, D_xz, D_yz) (``_precompute_operators_3D.py:17-61``)."""
    r = rearrange_indices_ext_int_3D(p)
    D = differentiation_matrix_1D(chebyshev_points(p)) / half_side_len
    I = np.eye(p)
    dx = np.kron(D, np.kron(I, I))[r][:, r]
    dy = np.kron(I, np.kron(D, I))[r][:, r]
    dz = np.kron(I, np.kron(I, D))[r][:, r]
    return dx, dy, dz, dx @ dx, dy @ dy, dz @ dz, dx @ dy, dx @ dz, dy @ dz


def get_face_1_idxes(p: int) -> np.ndarray:
    """x- face, (y, z) ij order (:220-224)."""
    return np.arange(p * p)


def get_face_2_idxes(p: int) -> np.ndarray:
    """x+ face (:227-231)."""
    return np.arange(p * p) + p * p


def get_face_3_idxes(p: int) -> np.ndarray:
    """y- face, (x, z) ij order (:234-246)."""
    return np.concatenate(
        [np.arange(p), np.arange(2 * p * p, 3 * p * p - 2 * p), np.arange(p * p, p * p + p)]
    )


def get_face_4_idxes(p: int) -> np.ndarray:
    """y+ face (:249-261)."""
    return np.concatenate(
        [
            np.arange(p * (p - 1), p * p),
            np.arange(3 * p * p - 2 * p, 4 * p * p - 4 * p),
            np.arange(2 * p * p - p, 2 * p * p),
        ]
    )


def _face_z(p: int, first_col, last_col, first_row, last_row, int_idxes) -> np.ndarray:
    out = np.zeros(p * p, dtype=np.int64)
    out[:p] = first_col
    out[-p:] = last_col
    c = p
    for i in range(p - 2):
        out[c] = first_row[i]
        c += 1
        out[c : c + p - 2] = int_idxes[i * (p - 2) : (i + 1) * (p - 2)]
        c += p - 2
        out[c] = last_row[i]
        c += 1
    return out


def get_face_5_idxes(p: int) -> np.ndarray:
    """z- face, (x, y) ij order (:264-293)."""
    return _face_z(
        p,
        np.arange(0, p * p, p),
        np.arange(p * p, 2 * p * p, p),
        np.arange(2 * p * p, 3 * p * p - 2 * p, p),
        np.arange(3 * p * p - 2 * p, 4 * p * p - 4 * p, p),
        np.arange(4 * p * p - 4 * p, 4 * p * p - 4 - p + (p - 2) ** 2),
    )


def get_face_6_idxes(p: int) -> np.ndarray:
    """z+ face (:296-320)."""
    u = p**3 - (p - 2) ** 3
    l = u - (p - 2) ** 2
    return _face_z(
        p,
        np.arange(p - 1, p * p, p),
        np.arange(p * p + p - 1, 2 * p * p, p),
        np.arange(2 * p * p + p - 1, 3 * p * p - 2 * p, p),
        np.arange(3 * p * p - 2 * p + p - 1, 4 * p * p - 4 * p, p),
        np.arange(l, u),
    )


FACE_IDX_FNS = (
    get_face_1_idxes,
    get_face_2_idxes,
    get_face_3_idxes,
    get_face_4_idxes,
    get_face_5_idxes,
    get_face_6_idxes,
)


def precompute_P_3D_DtN(p: int, q: int) -> np.ndarray:
    """Gauss (6 q^2) -> Chebyshev boundary (p^3 - (p-2)^3) with corner rows
    averaged over 3 faces and edge rows over 2 (:64-147)."""
    g = gauss_points(q)
    c = chebyshev_points(p)
    P = barycentric_lagrange_interpolation_matrix_2D(g, g, c, c)
    nb = p**3 - (p - 2) ** 3
    out = np.zeros((nb, 6 * q * q))
    for f, fn in enumerate(FACE_IDX_FNS):
        out[fn(p), f * q * q : (f + 1) * q * q] = P
    p2 = p * p
    corners = np.array([0, p - 1, p2 - p, p2 - 1, p2, p2 + p - 1, 2 * p2 - p, 2 * p2 - 1])
    out[corners] /= 3.0
    edges = np.concatenate(
        [
            np.arange(1, p - 1),
            np.arange(p2 + 1, p2 + p - 1),
            np.arange(p2 - p + 1, p2 - 1),
            np.arange(2 * p2 - p + 1, 2 * p2 - 1),
            np.arange(1, p - 1) * p,
            np.arange(p, p2 - p, p) + p2,
            np.arange(2 * p2, 3 * p2 - 2 * p, p),
            np.arange(3 * p2 - 2 * p, 4 * p2 - 4 * p, p),
            np.arange(2 * p - 1, p2 - p, p),
            np.arange(p2 + 2 * p - 1, 2 * p2 - p, p),
            np.arange(2 * p2 + p - 1, 3 * p2 - 2 * p, p),
            np.arange(3 * p2 - 2 * p + p - 1, 4 * p2 - 4 * p, p),
        ]
    )
    out[edges] /= 2.0
    return out


def precompute_Q_3D_DtN(p: int, q: int, du_dx, du_dy, du_dz) -> np.ndarray:
    """Leaf solution -> outward normal derivative on the 6 q^2 Gauss points (:150-210)."""
    p2 = p * p
    N = np.empty((6 * p2, p**3))
    N[:p2] = -du_dx[get_face_1_idxes(p)]
    N[p2 : 2 * p2] = du_dx[get_face_2_idxes(p)]
    N[2 * p2 : 3 * p2] = -du_dy[get_face_3_idxes(p)]
    N[3 * p2 : 4 * p2] = du_dy[get_face_4_idxes(p)]
    N[4 * p2 : 5 * p2] = -du_dz[get_face_5_idxes(p)]
    N[5 * p2 :] = du_dz[get_face_6_idxes(p)]
    g = gauss_points(q)
    c = chebyshev_points(p)
    Q = barycentric_lagrange_interpolation_matrix_2D(c, c, g, g)
    return np.kron(np.eye(6), Q) @ N
