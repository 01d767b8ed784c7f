"""ORACLE (test infrastructure only) -- exterior-scattering coupling.

Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` may import this package; it is the checker, never the product.

Restates ``examples/wave_scattering_utils.py:31-440`` in numpy: the ItI -> DtN
map, incident plane waves and their outward normal derivatives on the root
boundary (sides S, E, N, W), the boundary-integral system of the exterior
problem, the scattered field's incoming impedance data, and the whole
scattering solve through the FP64 HPS restatement (``oracle.hps2d``).

Pinned by (tests): the ItI -> DtN map inverting the Cayley transform
R = (T - i eta)(T + i eta)^-1 of a known T; the normal derivatives against
central finite differences of the plane wave along the outward normals; the
DtN of a q = 0 Helmholtz problem reproducing a plane wave's normal
derivative; the closed form u_scat = T^-1 du_in/dn - u_in when D = I/2; and a
vanishing scattered field for q = 0.
"""

from __future__ import annotations

import numpy as np

from . import hps2d
from . import spectral as osp


def dtn_from_iti(R, eta):
    """T = -i eta (R - I)^-1 (R + I) (``wave_scattering_utils.py:31-49``)."""
    n = R.shape[0]
    eye = np.eye(n)
    return -1j * eta * np.linalg.solve(R - eye, R + eye)


def plane_wave(k, pts, dirs):
    """exp(i k <x, (cos th, sin th)>), shape pts.shape[:-1] + (nsrc,)
    (``wave_scattering_utils.py:186-197``)."""
    d = np.stack([np.cos(dirs), np.sin(dirs)], axis=-1)  # (nsrc, 2)
    return np.exp(1j * k * (pts @ d.T))


def plane_wave_and_normals(k, bdry_pts, dirs):
    """(u_in, du_in/dn) on the root boundary, sides S, E, N, W of n/4 points
    with outward normals -y, +x, +y, -x (``wave_scattering_utils.py:139-183``)."""
    n = bdry_pts.shape[0]
    m = n // 4
    u = plane_wave(k, bdry_pts, dirs)
    c, s = np.cos(dirs)[None, :], np.sin(dirs)[None, :]
    side = np.minimum(np.arange(n) // m, 3)[:, None]
    kn = np.where(side == 0, -k * s, np.where(side == 1, k * c, np.where(side == 2, k * s, -k * c)))
    return u, 1j * kn * u


def scattering_system(S, D, T, bdry_pts, k, dirs):
    """A = I/2 - D + S T, b = S (du_in/dn - T u_in)
    (``wave_scattering_utils.py:96-136``)."""
    u, dn = plane_wave_and_normals(k, bdry_pts, dirs)
    A = 0.5 * np.eye(T.shape[0]) - D + S @ T
    return A, S @ (dn - T @ u)


def uscat_impedance(S, D, T, dirs, bdry_pts, k, eta):
    """Incoming impedance data of the scattered field
    (``wave_scattering_utils.py:200-244``)."""
    A, b = scattering_system(S, D, T, bdry_pts, k, dirs)
    u, dn = plane_wave_and_normals(k, bdry_pts, dirs)
    us = np.linalg.solve(A, b)
    return T @ (us + u) - dn + 1j * eta * us


def solve_scattering(bounds, L, p, k, q_fn, dirs, S, D):
    """Scattered field on the HPS grid (n_leaves, p^2, nsrc) for
    Delta u + k^2 (1 + q) u = 0 with incident plane waves, through the ItI
    leaf solve, merges, exterior coupling and down pass
    (``wave_scattering_utils.py:247-412``, before the interpolation)."""
    q = p - 2
    pts = osp.leaf_points_2D(bounds, L, p)
    half = (bounds[1] - bounds[0]) / 2 ** (L + 1)
    ops = osp.leaf_ops_2D(p, q, half, eta=k)
    qv = q_fn(pts)
    ones = np.ones(pts.shape[:2])
    coeffs = {"D_xx": ones, "D_yy": ones, "I": k**2 * (1 + qv)}
    A = hps2d.assemble_diff_operator(coeffs, ops)
    src = -(k**2) * qv[..., None] * plane_wave(k, pts, dirs)
    Tl, Y, h, v = hps2d.leaf_iti(A, src, ops["P_iti"], ops["QH"], ops["G"])
    S_lst, g_lst, R, _ = hps2d.merge_stage(Tl, h, L, "iti")
    T = dtn_from_iti(R, k)
    bdry = osp.boundary_gauss_points_2D(bounds, L, q)
    imp = uscat_impedance(S, D, T, dirs, bdry, k, k)
    return hps2d.down_pass(imp, S_lst, g_lst, Y, v, "iti")[0]
