"""Oracle solutions of the benchmark workloads (``jaxhps_amd.workloads``) at
sizes the CPU finishes in seconds: the stage-by-stage FP64 restatement
(``oracle.hps2d``) and the exact discrete solution of the same global system
(``oracle.global_sparse``)."""

import numpy as np

from oracle import global_sparse as gsp
from oracle import hps2d
from oracle import spectral as osp

_NAMES = {"D_xx_coefficients": "D_xx", "D_yy_coefficients": "D_yy", "D_xy_coefficients": "D_xy",
          "D_x_coefficients": "D_x", "D_y_coefficients": "D_y", "I_coefficients": "I"}


def _setup(dom, kw):
    L, p, q = dom.L, dom.p, dom.q
    r = dom.root
    bounds = (r.xmin, r.xmax, r.ymin, r.ymax)
    half = (r.xmax - r.xmin) / 2 ** (L + 1)
    iti = kw.get("use_ItI", False)
    ops = osp.leaf_ops_2D(p, q, half, eta=kw.get("eta") if iti else None)
    coeffs = {_NAMES[k]: np.asarray(v) for k, v in kw.items() if k in _NAMES and v is not None}
    A = hps2d.assemble_diff_operator(coeffs, ops)
    src = np.asarray(kw["source"])
    src = src if src.ndim == 3 else src[..., None]
    if iti:
        src = src.astype(np.complex128)
    return bounds, ops, A, src, iti


def stage_solution(dom, kw, g):
    """u (n_leaves, n, nsrc) of the FP64 stage-by-stage oracle."""
    bounds, ops, A, src, iti = _setup(dom, kw)
    gg = np.asarray(g)
    gg = gg if gg.ndim == 2 else gg[:, None]
    if iti:
        T, Y, h, v = hps2d.leaf_iti(A, src, ops["P_iti"], ops["QH"], ops["G"])
        S, gl, _, _ = hps2d.merge_stage(T, h, dom.L, "iti")
        return hps2d.down_pass(gg, S, gl, Y, v, "iti")[0]
    Y, T, v, h = hps2d.leaf_dtn(A, src, ops["P"], ops["Q"])
    S, gl, _, _ = hps2d.merge_stage(T, h, dom.L, "dtn")
    return hps2d.down_pass(gg, S, gl, Y, v, "dtn")[0]


def exact_solution(dom, kw, g):
    """u (n_leaves, n, nsrc): the global system solved by a sparse LU with
    extended-precision refinement."""
    bounds, ops, A, src, iti = _setup(dom, kw)
    gg = np.asarray(g)
    gg = gg if gg.ndim == 2 else gg[:, None]
    bp = dom.boundary_points
    if iti:
        return gsp.solve_iti_sparse(A, src, ops["P_iti"], ops["QH"], ops["G"], bounds, dom.L, dom.q, gg, bp)
    return gsp.solve_dtn_sparse(A, src, ops["P"], ops["Q"], bounds, dom.L, dom.q, gg, bp)


def rel2(a, b):
    a = np.asarray(a).reshape(np.asarray(b).shape)
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(np.asarray(b).ravel()))
