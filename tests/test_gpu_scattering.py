"""GPU parity of the exterior-scattering coupling (ItI -> DtN, plane waves,
boundary-integral system, impedance data, the whole scattering solve)
against the oracle and its closed forms."""

import numpy as np
import pytest
import torch

from oracle import interp as oi
from oracle import scattering as osc
from oracle import spectral as osp

pytestmark = pytest.mark.gpu

BOUNDS = (-1.0, 1.0, -1.0, 1.0)


def _np(t):
    return t.detach().cpu().numpy()


def _rand(rng, n, m):
    return rng.standard_normal((n, m)) + 1j * rng.standard_normal((n, m))


def _rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


@pytest.mark.parametrize("n", [112, 448, 1792])
def test_dtn_from_iti_recovers_T_from_its_cayley_transform(n):
    from jaxhps_amd.scattering import get_DtN_from_ItI

    print(f"iti_to_dtn n={n}", flush=True)
    rng = np.random.default_rng(n)
    eta = 6.0
    T = _rand(rng, n, n) / np.sqrt(n) * eta
    eye = np.eye(n)
    R = (T - 1j * eta * eye) @ np.linalg.inv(T + 1j * eta * eye)
    got = _np(get_DtN_from_ItI(torch.as_tensor(R, device="cuda"), eta))
    assert _rel(got, osc.dtn_from_iti(R, eta)) < 1e-10
    assert _rel(got, T) < 1e-9


def test_plane_waves_match_oracle():
    from jaxhps_amd.scattering import get_uin, get_uin_and_normals

    k, dirs = 9.0, np.array([0.0, 0.4, 2.2, -2.9])
    bdry = osp.boundary_gauss_points_2D(BOUNDS, 2, 6)
    u, dn = get_uin_and_normals(k, bdry, dirs)
    wu, wdn = osc.plane_wave_and_normals(k, bdry, dirs)
    assert np.abs(_np(u) - wu).max() < 1e-13 and np.abs(_np(dn) - wdn).max() < 1e-12
    pts = osp.leaf_points_2D(BOUNDS, 2, 8)
    got = _np(get_uin(k, pts, dirs))
    assert got.shape == (16, 64, 4)
    assert np.abs(got - osc.plane_wave(k, pts, dirs)).max() < 1e-13


def _sdt(rng, n):
    T = _rand(rng, n, n) + 3 * n * np.eye(n)
    S = _rand(rng, n, n) / n
    D = _rand(rng, n, n) / n
    return S, D, T


@pytest.mark.parametrize("n,nsrc", [(64, 1), (448, 5)])
def test_scattering_system_and_impedance_match_oracle(n, nsrc):
    from jaxhps_amd.scattering import get_scattering_uscat_impedance, setup_scattering_lin_system

    rng = np.random.default_rng(n + nsrc)
    S, D, T = _sdt(rng, n)
    bdry = osp.boundary_gauss_points_2D(BOUNDS, 2, n // 16)
    dirs = rng.uniform(-np.pi, np.pi, nsrc)
    k, eta = 7.0, 7.0
    A, b = setup_scattering_lin_system(S, D, T, bdry, k, dirs)
    wA, wb = osc.scattering_system(S, D, T, bdry, k, dirs)
    assert _rel(_np(A), wA) < 1e-12 and _rel(_np(b), wb) < 1e-12
    imp = _np(get_scattering_uscat_impedance(S, D, T, dirs, bdry, k, eta))
    assert imp.shape == (n, nsrc)
    assert _rel(imp, osc.uscat_impedance(S, D, T, dirs, bdry, k, eta)) < 1e-10
    # D = I/2: imp = i eta (T^-1 du_in/dn - u_in).  The identity holds to the
    # rounding of the boundary-integral system (condition ~6e2 at n = 448):
    # LAPACK through the oracle lands at 5e-11 from the closed form, so the
    # bound is a few times that
    u, dn = osc.plane_wave_and_normals(k, bdry, dirs)
    closed = 1j * eta * (np.linalg.solve(T, dn) - u)
    imp2 = _np(get_scattering_uscat_impedance(S, 0.5 * np.eye(n), T, dirs, bdry, k, eta))
    lapack = _rel(osc.uscat_impedance(S, 0.5 * np.eye(n), T, dirs, bdry, k, eta), closed)
    assert _rel(imp2, closed) < max(1e-10, 4 * lapack)


def _gauss_bump(x):
    return 0.6 * np.exp(-((x[..., 0] - 0.1) ** 2 + (x[..., 1] + 0.2) ** 2) / 0.08)


def test_scattering_solve_matches_oracle():
    from jaxhps_amd.scattering import solve_scattering_problem

    p, L, k, ngrid = 12, 2, 6.0, 9
    n = 4 * (p - 2) * 2**L
    rng = np.random.default_rng(5)
    S = 0.3 * _rand(rng, n, n) / n
    D = 0.3 * _rand(rng, n, n) / n
    dirs = np.array([0.3, 1.7, -2.4])
    vals, target, secs = solve_scattering_problem(L, p, ngrid, k, _gauss_bump, BOUNDS, dirs,
                                                  torch.as_tensor(S, device="cuda"), torch.as_tensor(D, device="cuda"))
    u_hps = osc.solve_scattering(BOUNDS, L, p, k, _gauss_bump, dirs, S, D)
    xs = np.linspace(BOUNDS[0], BOUNDS[1], ngrid)
    want, wtgt = oi.interp_from_interior_2D(BOUNDS, L, p, u_hps, xs, xs)
    assert tuple(vals.shape) == (ngrid, ngrid, 3)
    assert np.array_equal(np.asarray(target), wtgt)
    assert _rel(_np(vals), want) < 1e-9
    assert secs > 0


def test_no_scatterer_no_scattered_field():
    from jaxhps_amd.scattering import solve_scattering_problem

    p, L, k = 16, 2, 5.0
    n = 4 * (p - 2) * 2**L
    rng = np.random.default_rng(6)
    S = torch.as_tensor(0.3 * _rand(rng, n, n) / n, device="cuda")
    D = torch.as_tensor(0.3 * _rand(rng, n, n) / n, device="cuda")
    vals, _, _ = solve_scattering_problem(L, p, 7, k, lambda x: np.zeros(x.shape[:-1]), BOUNDS, np.array([0.5, 2.0]),
                                          S, D)
    assert np.abs(_np(vals)).max() < 1e-8
