"""The exterior-scattering oracle pinned by closed forms: the Cayley map
between DtN and ItI, finite differences of the plane wave, the DtN of a
q = 0 Helmholtz problem, a D = I/2 closed form and q = 0 => no scattering."""

import numpy as np

from oracle import hps2d
from oracle import scattering as sc
from oracle import spectral as osp

BOUNDS = (-1.0, 1.0, -1.0, 1.0)


def _rand(rng, n, m):
    return rng.standard_normal((n, m)) + 1j * rng.standard_normal((n, m))


def test_dtn_from_iti_inverts_the_cayley_map():
    rng = np.random.default_rng(0)
    n, eta = 24, 3.5
    T = _rand(rng, n, n)
    R = (T - 1j * eta * np.eye(n)) @ np.linalg.inv(T + 1j * eta * np.eye(n))
    assert np.abs(sc.dtn_from_iti(R, eta) - T).max() < 1e-11 * np.abs(T).max()


def test_normals_match_finite_differences():
    k, dirs = 7.0, np.array([0.3, 2.0, -1.1])
    bdry = osp.boundary_gauss_points_2D(BOUNDS, 1, 5)
    n = bdry.shape[0]
    normals = np.repeat(np.array([[0, -1.0], [1.0, 0], [0, 1.0], [-1.0, 0]]), n // 4, axis=0)
    u, dn = sc.plane_wave_and_normals(k, bdry, dirs)
    e = 1e-5
    fd = (sc.plane_wave(k, bdry + e * normals, dirs) - sc.plane_wave(k, bdry - e * normals, dirs)) / (2 * e)
    assert np.abs(u - sc.plane_wave(k, bdry, dirs)).max() == 0
    assert np.abs(dn - fd).max() < 1e-7 * k


def _helmholtz_q0(p, L, k):
    q = p - 2
    half = (BOUNDS[1] - BOUNDS[0]) / 2 ** (L + 1)
    ops = osp.leaf_ops_2D(p, q, half, eta=k)
    pts = osp.leaf_points_2D(BOUNDS, L, p)
    ones = np.ones(pts.shape[:2])
    A = hps2d.assemble_diff_operator({"D_xx": ones, "D_yy": ones, "I": k**2 * ones}, ops)
    T, Y, h, v = hps2d.leaf_iti(A, np.zeros(pts.shape[:2] + (1,)), ops["P_iti"], ops["QH"], ops["G"])
    R = hps2d.merge_stage(T, h, L, "iti")[2]
    return sc.dtn_from_iti(R, k), osp.boundary_gauss_points_2D(BOUNDS, L, q)


def test_dtn_of_helmholtz_maps_plane_waves_to_their_normal_derivatives():
    k = 4.0
    T, bdry = _helmholtz_q0(16, 1, k)
    u, dn = sc.plane_wave_and_normals(k, bdry, np.array([0.0, 0.7, 2.5]))
    # discretisation error: 4e-3 at p = 8, 2e-6 at p = 12, 1e-10 at p = 16
    assert np.abs(T @ u - dn).max() < 1e-9 * np.abs(dn).max()


def test_impedance_closed_form_when_D_is_half_identity():
    """A = S T when D = I/2, so u_scat = T^-1 du_in/dn - u_in,
    du_scat/dn = 0 and imp = i eta u_scat."""
    rng = np.random.default_rng(1)
    k, eta = 5.0, 5.0
    bdry = osp.boundary_gauss_points_2D(BOUNDS, 1, 4)
    n = bdry.shape[0]
    T = _rand(rng, n, n) + 4 * n * np.eye(n)
    S = _rand(rng, n, n) + n * np.eye(n)
    dirs = np.array([0.2, 1.9])
    u, dn = sc.plane_wave_and_normals(k, bdry, dirs)
    want = 1j * eta * (np.linalg.solve(T, dn) - u)
    got = sc.uscat_impedance(S, 0.5 * np.eye(n), T, dirs, bdry, k, eta)
    assert np.abs(got - want).max() < 1e-10 * np.abs(want).max()


def test_no_scatterer_gives_no_scattered_field():
    rng = np.random.default_rng(2)
    k, p, L = 4.0, 16, 1
    n = 4 * (p - 2) * 2**L
    S = 0.01 * _rand(rng, n, n) / n
    D = 0.05 * _rand(rng, n, n) / n
    u = sc.solve_scattering(BOUNDS, L, p, k, lambda x: np.zeros(x.shape[:-1]), np.array([0.4, 1.3]), S, D)
    assert u.shape == (4**L, p * p, 2)
    assert np.abs(u).max() < 1e-8
