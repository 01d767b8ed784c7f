"""Closed-form 2D test problems, restated in numpy from the reference's
``tests/test_accuracy/cases.py`` (domain [-pi/2, pi/2]^2, eta = 1)."""

import numpy as np

XMIN, XMAX, YMIN, YMAX = -np.pi / 2, np.pi / 2, -np.pi / 2, np.pi / 2
BOUNDS = (XMIN, XMAX, YMIN, YMAX)
ETA = 1.0
A2 = (np.pi / 2) ** 2


def poly_dirichlet(x):
    """x^2 - y^2 (harmonic) -- ``cases.py:61-64``."""
    return x[..., 0] ** 2 - x[..., 1] ** 2


def poly_dirichlet_dx(x):
    return 2 * x[..., 0]


def poly_dirichlet_dy(x):
    return -2 * x[..., 1]


def poly_part_soln(x):
    """y (x^2 - a)(y^2 - a), zero on the boundary -- ``cases.py:79-85``."""
    return x[..., 1] * (x[..., 1] ** 2 - A2) * (x[..., 0] ** 2 - A2)


def poly_source(x):
    """Laplacian of poly_part_soln -- ``cases.py:88-92``."""
    return 2 * x[..., 1] * (x[..., 1] ** 2 - A2) + 6 * x[..., 1] * (x[..., 0] ** 2 - A2)


def normal_derivative(pts, dudx, dudy):
    """Outward normal derivative on the root boundary (S, E, N, W blocks)."""
    n4 = pts.shape[0] // 4
    return np.concatenate(
        [
            -dudy(pts[:n4]),
            dudx(pts[n4 : 2 * n4]),
            dudy(pts[2 * n4 : 3 * n4]),
            -dudx(pts[3 * n4 :]),
        ]
    )


# Complex-coefficient ItI case (``cases.py:220-260``): Lap u + c u = f
CPLX = 1.0 + 0.1j


def cc_u(x):
    return x[..., 0] ** 3 + 3 * x[..., 1] ** 2


def cc_source(x):
    return 6 * x[..., 0] + 6 + CPLX * cc_u(x)


def cc_dx(x):
    return 3 * x[..., 0] ** 2


def cc_dy(x):
    return 6 * x[..., 1]


# Helmholtz-like case (``cases.py:149-209``): Lap u + q u = f with
# q = 1 + exp(-|x|^2), u = exp(i pi x) + exp(i pi y)
def hz_q(x):
    return 1.0 + np.exp(-x[..., 0] ** 2 - x[..., 1] ** 2)


def hz_u(x):
    return np.exp(1j * np.pi * x[..., 0]) + np.exp(1j * np.pi * x[..., 1])


def hz_f(x):
    return -(np.pi**2) * hz_u(x) + hz_q(x) * hz_u(x)


def hz_dx(x):
    return 1j * np.pi * np.exp(1j * np.pi * x[..., 0])


def hz_dy(x):
    return 1j * np.pi * np.exp(1j * np.pi * x[..., 1])
