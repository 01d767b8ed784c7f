"""Seeded synthetic problems of the benchmark configurations.

All on [-1, 1]^d with q = p - 2, numpy.random.default_rng(seed).  Each
function returns (Domain, PDEProblem kwargs, boundary data) -- host numpy
arrays -- so the same inputs feed the GPU path and the CPU oracle.
"""

from __future__ import annotations

import numpy as np

from ._discretization_tree import DiscretizationNode2D, DiscretizationNode3D
from ._domain import Domain


def gaussian_bumps(x, centers, widths, amps):
    """sum_i amps_i exp(-|x - c_i|^2 / w_i^2) and its gradient."""
    val = np.zeros(x.shape[:-1])
    grad = np.zeros(x.shape)
    for c, w, a in zip(centers, widths, amps):
        d = x - c
        e = a * np.exp(-np.sum(d * d, axis=-1) / w**2)
        val += e
        grad += (-2.0 / w**2) * e[..., None] * d
    return val, grad


def config_A(L: int = 3, p: int = 8):
    """2D Poisson -Lap u = f on [-1, 1]^2, DtN, with the manufactured solution
    u = sin(pi x) sin(pi y): D_xx = D_yy = -1, f = 2 pi^2 u, Dirichlet data u
    on the root boundary's Gauss points."""
    root = DiscretizationNode2D(-1.0, 1.0, -1.0, 1.0)
    dom = Domain(p=p, q=p - 2, root=root, L=L)
    x = dom.interior_points
    ones = np.ones(x.shape[:2])
    u = np.sin(np.pi * x[..., 0]) * np.sin(np.pi * x[..., 1])
    kwargs = dict(D_xx_coefficients=-ones, D_yy_coefficients=-ones, source=2 * np.pi**2 * u)
    bp = dom.boundary_points
    g = np.sin(np.pi * bp[:, 0]) * np.sin(np.pi * bp[:, 1])
    return dom, kwargs, g


def exact_A(dom):
    """The manufactured solution of config_A at the leaf points."""
    x = dom.interior_points
    return np.sin(np.pi * x[..., 0]) * np.sin(np.pi * x[..., 1])


def config_B(L: int = 6, p: int = 16, nsrc: int = 16, seed: int = 0):
    """2D variable-coefficient elliptic -div(a grad u) + c u = f, DtN.

    a = 1 + 0.5 * (8 Gaussian bumps, centers U[-0.8,0.8]^2, widths U[0.05,0.2]);
    c = 4 bumps scaled into [0, 1]; nsrc random Dirichlet boundary data ~ N(0,1),
    zero source of shape (n_leaves, p^2, nsrc).
    -div(a grad u) = -a Lap u - grad a . grad u, so D_xx = D_yy = -a,
    D_x = -da/dx, D_y = -da/dy, I = c.
    """
    rng = np.random.default_rng(seed)
    root = DiscretizationNode2D(-1.0, 1.0, -1.0, 1.0)
    dom = Domain(p=p, q=p - 2, root=root, L=L)
    x = dom.interior_points
    cen = rng.uniform(-0.8, 0.8, (8, 2))
    wid = rng.uniform(0.05, 0.2, 8)
    bump, gb = gaussian_bumps(x, cen, wid, np.ones(8))
    a = 1.0 + 0.5 * bump
    ga = 0.5 * gb
    cc = rng.uniform(-0.8, 0.8, (4, 2))
    cw = rng.uniform(0.2, 0.6, 4)
    cb, _ = gaussian_bumps(x, cc, cw, np.ones(4))
    c = cb / max(cb.max(), 1e-300)
    kwargs = dict(
        D_xx_coefficients=-a,
        D_yy_coefficients=-a,
        D_x_coefficients=-ga[..., 0],
        D_y_coefficients=-ga[..., 1],
        I_coefficients=c,
        source=np.zeros(x.shape[:2] + (nsrc,)),
    )
    g = rng.standard_normal((dom.boundary_points.shape[0], nsrc))
    return dom, kwargs, g


def plane_wave_impedance(pts, k, eta, alpha):
    """Incoming impedance u_n + i eta u of u = exp(i k (x cos a + y sin a)) on
    the root boundary (S, E, N, W blocks, outward normals)."""
    d = np.array([np.cos(alpha), np.sin(alpha)])
    u = np.exp(1j * k * (pts @ d))
    n4 = pts.shape[0] // 4
    normals = np.concatenate(
        [np.tile([0.0, -1.0], (n4, 1)), np.tile([1.0, 0.0], (n4, 1)), np.tile([0.0, 1.0], (n4, 1)), np.tile([-1.0, 0.0], (n4, 1))]
    )
    un = 1j * k * (normals @ d) * u
    return un + 1j * eta * u


def config_C(L: int = 7, p: int = 16, k: float = 100.0, seed: int = 0):
    """2D interior Helmholtz Lap u + k^2 (1 - q) u = 0 with ItI merges, eta = k.

    q = 20 Gaussian bumps (amplitudes U[0.5,1.5], widths U[0.02,0.05], centers
    U[-0.9,0.9]^2); impedance data of a plane wave at a random angle."""
    rng = np.random.default_rng(seed)
    root = DiscretizationNode2D(-1.0, 1.0, -1.0, 1.0)
    dom = Domain(p=p, q=p - 2, root=root, L=L)
    x = dom.interior_points
    cen = rng.uniform(-0.9, 0.9, (20, 2))
    wid = rng.uniform(0.02, 0.05, 20)
    amp = rng.uniform(0.5, 1.5, 20)
    q, _ = gaussian_bumps(x, cen, wid, amp)
    ones = np.ones(x.shape[:2])
    kwargs = dict(
        D_xx_coefficients=ones,
        D_yy_coefficients=ones,
        I_coefficients=k**2 * (1.0 - q),
        source=np.zeros(x.shape[:2], dtype=np.complex128),
        use_ItI=True,
        eta=k,
    )
    alpha = rng.uniform(0.0, 2 * np.pi)
    g = plane_wave_impedance(dom.boundary_points, k, k, alpha)
    return dom, kwargs, g


def config_D(L: int = 3, p: int = 10, nsrc: int = 4, seed: int = 0):
    """3D variable-coefficient Poisson on [-1,1]^3: -div(a grad u) = f with
    a = 1 + 0.5 * (6 bumps, widths U[0.1,0.3]); manufactured
    u_k = sin(pi x) sin(pi y) sin(pi z) * (1 + 0.25 k (x + y + z)) for nsrc
    right-hand sides (f from the analytic derivatives), Dirichlet data u_k."""
    rng = np.random.default_rng(seed)
    root = DiscretizationNode3D(-1.0, 1.0, -1.0, 1.0, -1.0, 1.0)
    dom = Domain(p=p, q=p - 2, root=root, L=L)
    x = dom.interior_points
    cen = rng.uniform(-0.8, 0.8, (6, 3))
    wid = rng.uniform(0.1, 0.3, 6)
    bump, gb = gaussian_bumps(x, cen, wid, np.ones(6))
    a = 1.0 + 0.5 * bump
    ga = 0.5 * gb

    def u_and_derivs(pts, kk):
        X, Y, Z = pts[..., 0], pts[..., 1], pts[..., 2]
        s = np.sin(np.pi * X) * np.sin(np.pi * Y) * np.sin(np.pi * Z)
        sx = np.pi * np.cos(np.pi * X) * np.sin(np.pi * Y) * np.sin(np.pi * Z)
        sy = np.pi * np.sin(np.pi * X) * np.cos(np.pi * Y) * np.sin(np.pi * Z)
        sz = np.pi * np.sin(np.pi * X) * np.sin(np.pi * Y) * np.cos(np.pi * Z)
        m = 1 + 0.25 * kk * (X + Y + Z)
        u = s * m
        grad = np.stack([sx * m + 0.25 * kk * s, sy * m + 0.25 * kk * s, sz * m + 0.25 * kk * s], -1)
        # Laplacian: m Lap s + 2 grad s . grad m (Lap m = 0)
        lap = -3 * np.pi**2 * s * m + 2 * 0.25 * kk * (sx + sy + sz)
        return u, grad, lap

    srcs, bds = [], []
    for kk in range(nsrc):
        u, gu, lap = u_and_derivs(x, kk)
        f = -(a * lap + np.sum(ga * gu, axis=-1))
        srcs.append(f)
        ub, _, _ = u_and_derivs(dom.boundary_points, kk)
        bds.append(ub)
    kwargs = dict(
        D_xx_coefficients=-a,
        D_yy_coefficients=-a,
        D_zz_coefficients=-a,
        D_x_coefficients=-ga[..., 0],
        D_y_coefficients=-ga[..., 1],
        D_z_coefficients=-ga[..., 2],
        source=np.stack(srcs, -1),
    )
    return dom, kwargs, np.stack(bds, -1)


def exact_D(dom, nsrc: int = 4):
    """Manufactured solutions of config_D at the leaf points, (n_leaves, p^3, nsrc)."""
    x = dom.interior_points
    out = []
    for kk in range(nsrc):
        X, Y, Z = x[..., 0], x[..., 1], x[..., 2]
        out.append(np.sin(np.pi * X) * np.sin(np.pi * Y) * np.sin(np.pi * Z) * (1 + 0.25 * kk * (X + Y + Z)))
    return np.stack(out, -1)


def sine_freqs(k: float, n_terms: int = 25) -> np.ndarray:
    """The first n_terms frequency pairs (k1, k2), 1 <= k1, k2, |(k1, k2)| <= 2k,
    in increasing norm (ties in (k1, k2) order) -- get_freqs_up_to_2k of
    ``examples/sine_transform.py:103-140``."""
    gamma = int(k * 2)
    f = np.array([[k1, k2] for k1 in range(1, gamma + 1) for k2 in range(1, gamma + 1)
                  if np.sqrt(k1**2 + k2**2) <= gamma], dtype=np.float64)
    order = np.argsort(np.linalg.norm(f, axis=1), kind="stable")
    return f[order][:n_terms]


def sine_basis(freqs: np.ndarray, pts: np.ndarray) -> np.ndarray:
    """(n_terms, n_leaves, n): sin(pi/2 k1 (x+1)) sin(pi/2 k2 (y+1)), the basis of
    adjoint_nu_sinetransform (``examples/sine_transform.py:61-100``)."""
    sx = np.sin(np.pi / 2 * freqs[:, 0, None, None] * (pts[None, ..., 0] + 1.0))
    sy = np.sin(np.pi / 2 * freqs[:, 1, None, None] * (pts[None, ..., 1] + 1.0))
    return sx * sy


def config_E(L: int = 6, p: int = 16, k: float = 40.0, n_theta: int = 8, n_terms: int = 25, seed: int = 0):
    """Inverse-scattering gradient workload: Lap u + k^2 (1 - q_theta) u = 0,
    ItI with eta = k, q_theta = sum_j theta_j basis_j over the first 25 sine
    frequencies, plane-wave impedance data; thetas ~ N(0, 0.1^2) (n_theta of
    them) and a target theta_star.  Returns (Domain, PDEProblem kwargs without
    I_coefficients, boundary data, basis (n_terms, n_leaves, n), thetas,
    theta_star)."""
    rng = np.random.default_rng(seed)
    root = DiscretizationNode2D(-1.0, 1.0, -1.0, 1.0)
    dom = Domain(p=p, q=p - 2, root=root, L=L)
    x = dom.interior_points
    basis = sine_basis(sine_freqs(k, n_terms), x)
    ones = np.ones(x.shape[:2])
    kwargs = dict(D_xx_coefficients=ones, D_yy_coefficients=ones, source=np.zeros(x.shape[:2], dtype=np.complex128),
                  use_ItI=True, eta=k)
    alpha = rng.uniform(0.0, 2 * np.pi)
    g = plane_wave_impedance(dom.boundary_points, k, k, alpha)
    thetas = 0.1 * rng.standard_normal((n_theta, n_terms))
    theta_star = 0.1 * rng.standard_normal(n_terms)
    return dom, kwargs, g, basis, thetas, theta_star
