"""Exterior-scattering coupling of the top-level operator
(``examples/wave_scattering_utils.py:31-440``), on the device.

The interior problem Delta u + k^2 (1 + q) u = 0 is built with the ItI
solver; its root ItI map R becomes the DtN map T, which couples to the
exterior through single- and double-layer matrices S, D on the root boundary
(the boundary-integral equation A u_scat = b, A = I/2 - D + S T).  Its
solution gives the incoming impedance data of the scattered field for the
down pass.  All dense work (two n x n LU factorisations, n^3 products) runs in
the library's kernels (``hps_iti_to_dtn``, ``hps_scattering_system``,
``hps_scattering_impedance``, ``hps_plane_wave``); complex128 throughout.
"""

from __future__ import annotations

from timeit import default_timer

import numpy as np
import torch

from . import _lib
from ._lib import HPS_C128, check, load_library, ptr, raise_on_info, workspace

_C = torch.complex128


def _dev(*ts):
    for t in ts:
        if isinstance(t, torch.Tensor) and t.is_cuda:
            return t.device
    return torch.device("cuda", 0)


def _c128(x, dev):
    return torch.as_tensor(x).to(device=dev, dtype=_C).contiguous()


def _f64(x, dev):
    return torch.as_tensor(np.asarray(x, dtype=np.float64) if not isinstance(x, torch.Tensor) else x).to(
        device=dev, dtype=torch.float64).contiguous()


def _dirs(source_directions, dev):
    d = _f64(source_directions, dev)
    return d.reshape(-1)


def get_DtN_from_ItI(R, eta: float):
    """T = -i eta (R - I)^{-1} (R + I) (``wave_scattering_utils.py:31-49``)."""
    dev = _dev(R)
    Rd = _c128(R, dev)
    n = Rd.shape[0]
    if Rd.shape != (n, n):
        raise ValueError("R must be square")
    lib = load_library()
    h = _lib.handle()
    T = torch.empty_like(Rd)
    info = torch.zeros(1, dtype=torch.int32, device=dev)
    ws = workspace(lib.hps_iti_to_dtn_workspace(n), dev)
    rc = lib.hps_iti_to_dtn(h, HPS_C128, n, ptr(Rd), n, float(eta), ptr(T), n, ptr(info), ptr(ws), ws.numel())
    check(rc, h, "hps_iti_to_dtn")
    raise_on_info(info, "get_DtN_from_ItI: R - I")
    return T


def _plane_wave(k, pts, source_directions, n_per_side=0):
    dev = _dev(pts, source_directions)
    P = _f64(pts, dev)
    if P.shape[-1] != 2:
        raise ValueError("points must have a trailing axis of 2")
    d = _dirs(source_directions, dev)
    npts, nsrc = P.numel() // 2, d.numel()
    uin = torch.empty(P.shape[:-1] + (nsrc,), dtype=_C, device=dev)
    dn = torch.empty_like(uin) if n_per_side else None
    lib = load_library()
    h = _lib.handle()
    rc = lib.hps_plane_wave(h, npts, nsrc, ptr(P), ptr(d), float(k), int(n_per_side), ptr(uin), ptr(dn))
    check(rc, h, "hps_plane_wave")
    return uin, dn


def get_uin(k: float, pts, source_directions):
    """exp(i k <x, (cos th, sin th)>) with shape pts.shape[:-1] + (n_sources,)
    (``wave_scattering_utils.py:186-197``)."""
    return _plane_wave(k, pts, source_directions)[0]


def get_uin_and_normals(k: float, bdry_pts, source_directions):
    """(u_in, du_in/dn) on the root boundary points (sides S, E, N, W)
    (``wave_scattering_utils.py:139-183``)."""
    n = int(bdry_pts.shape[0])
    if n < 4:
        raise ValueError("need the four sides of the root boundary")
    return _plane_wave(k, bdry_pts, source_directions, n // 4)


def _check_sdt(S, D, T):
    n = T.shape[0]
    for nm, M in (("S", S), ("D", D), ("T", T)):
        if tuple(M.shape) != (n, n):
            raise ValueError(f"{nm} must have shape {(n, n)}")
    return n


def setup_scattering_lin_system(S, D, T_int, gauss_bdry_pts, k: float, source_directions):
    """A = I/2 - D + S T, b = S (du_in/dn - T u_in)
    (``wave_scattering_utils.py:96-136``)."""
    dev = _dev(S, D, T_int)
    S, D, T = (_c128(x, dev) for x in (S, D, T_int))
    n = _check_sdt(S, D, T)
    P = _f64(gauss_bdry_pts, dev)
    d = _dirs(source_directions, dev)
    nsrc = d.numel()
    A = torch.empty((n, n), dtype=_C, device=dev)
    b = torch.empty((n, nsrc), dtype=_C, device=dev)
    lib = load_library()
    h = _lib.handle()
    ws = workspace(lib.hps_scattering_workspace(n, nsrc), dev)
    rc = lib.hps_scattering_system(h, HPS_C128, n, nsrc, ptr(S), ptr(D), ptr(T), ptr(P), ptr(d), float(k), ptr(A),
                                   ptr(b), ptr(ws), ws.numel())
    check(rc, h, "hps_scattering_system")
    return A, b


def get_scattering_uscat_impedance(S, D, T, source_dirs, bdry_pts, k: float, eta: float):
    """Incoming impedance data (n, n_sources) of the scattered field
    (``wave_scattering_utils.py:200-244``)."""
    dev = _dev(S, D, T)
    S, D, T = (_c128(x, dev) for x in (S, D, T))
    n = _check_sdt(S, D, T)
    P = _f64(bdry_pts, dev)
    if tuple(P.shape) != (n, 2):
        raise ValueError(f"bdry_pts must have shape {(n, 2)}")
    d = _dirs(source_dirs, dev)
    nsrc = d.numel()
    imp = torch.empty((n, nsrc), dtype=_C, device=dev)
    info = torch.zeros(1, dtype=torch.int32, device=dev)
    lib = load_library()
    h = _lib.handle()
    ws = workspace(lib.hps_scattering_workspace(n, nsrc), dev)
    rc = lib.hps_scattering_impedance(h, HPS_C128, n, nsrc, ptr(S), ptr(D), ptr(T), ptr(P), ptr(d), float(k),
                                      float(eta), ptr(imp), ptr(info), ptr(ws), ws.numel())
    check(rc, h, "hps_scattering_impedance")
    raise_on_info(info, "get_scattering_uscat_impedance: boundary-integral system")
    return imp


def load_SD_matrices(fp: str):
    """S and D from a MATLAB v7.3 file (``wave_scattering_utils.py:52-93``):
    datasets "S" and "D" of (real, imag) records, stored column-major."""
    try:
        import h5py
    except ImportError as e:
        raise ImportError("load_SD_matrices reads MATLAB v7.3 (HDF5) files and needs h5py") from e
    with h5py.File(fp, "r") as f:
        if "S" not in f or "D" not in f:
            raise ValueError("File must contain both 'S' and 'D' matrices")

        def c(a):
            a = a[:].T
            return torch.as_tensor(np.asarray(a["real"] + 1j * a["imag"]))

        return c(f["S"]), c(f["D"])


def solve_scattering_problem(l: int, p: int, n: int, k: float, q_fn, domain_bounds, source_dirs, S, D):
    """Scattered field of incident plane waves off the potential q on an n x n
    grid (``wave_scattering_utils.py:247-440``): ItI build, ItI -> DtN,
    exterior coupling, down pass, interpolation to the grid.  Returns
    (u_scat (n, n, n_sources) on the device, target points, seconds)."""
    from ._build_solver import build_solver
    from ._discretization_tree import DiscretizationNode2D
    from ._domain import Domain
    from ._pdeproblem import PDEProblem
    from ._solve import solve

    xmin, xmax, ymin, ymax = (float(v) for v in domain_bounds)
    domain = Domain(p=p, q=p - 2, root=DiscretizationNode2D(xmin=xmin, xmax=xmax, ymin=ymin, ymax=ymax), L=l)
    pts = domain.interior_points
    qv = np.asarray(q_fn(pts))
    ones = np.ones(pts.shape[:2])
    dev = _dev(S, D)
    uin = get_uin(k, pts, source_dirs)
    source = -(k**2) * torch.as_tensor(qv, device=dev)[..., None] * uin
    pde = PDEProblem(domain=domain, D_xx_coefficients=ones, D_yy_coefficients=ones, I_coefficients=k**2 * (1 + qv),
                     source=source, use_ItI=True, eta=k)
    t0 = default_timer()
    R = build_solver(pde, return_top_T=True)
    T = get_DtN_from_ItI(R, pde.eta)
    imp = get_scattering_uscat_impedance(S, D, T, source_dirs, domain.boundary_points, k, k)
    del R, T
    u = solve(pde, imp)
    xs, ys = np.linspace(xmin, xmax, n), np.linspace(ymin, ymax, n)
    vals, target = domain.interp_from_interior_points(u, xs, ys)
    torch.cuda.synchronize(dev)
    return vals, target, default_timer() - t0


__all__ = [
    "get_DtN_from_ItI",
    "get_uin",
    "get_uin_and_normals",
    "setup_scattering_lin_system",
    "get_scattering_uscat_impedance",
    "load_SD_matrices",
    "solve_scattering_problem",
]
