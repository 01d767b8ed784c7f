"""``PDEProblem``: coefficients, precomputed leaf operators and solver state.

Signature, validation and attribute names follow ``_pdeproblem.py:25-234``
of the reference.  Coefficients and sources may be numpy arrays or torch
tensors; they are moved to the GPU by the build.  Leaf operators are built
once on the host in float64 (every uniform leaf has the same half side
length, ``_pdeproblem.py:129-138``) and cached on the device per dtype.
"""

from __future__ import annotations

from typing import List

import numpy as np
import torch

from ._discretization_tree import DiscretizationNode3D, get_all_leaves
from ._domain import Domain
from ._precompute_operators import (
    precompute_diff_operators_2D,
    precompute_diff_operators_3D,
    precompute_G_2D_ItI,
    precompute_N_matrix_2D,
    precompute_N_tilde_matrix_2D,
    precompute_P_2D_DtN,
    precompute_P_2D_ItI,
    precompute_P_3D_DtN,
    precompute_Q_2D_DtN,
    precompute_Q_3D_DtN,
    precompute_QH_2D_ItI,
)

COEFF_NAMES_2D = ("D_xx", "D_xy", "D_yy", "D_x", "D_y", "I")
COEFF_NAMES_3D = ("D_xx", "D_xy", "D_yy", "D_xz", "D_yz", "D_zz", "D_x", "D_y", "D_z", "I")


def _shape(a):
    return tuple(a.shape)


def check_input_shapes(source, use_ItI, expected_shape, **coeffs) -> None:
    """Every present coefficient must have shape (n_leaves, p^d) (``_pdeproblem.py:469-505``)."""
    for name, arr in coeffs.items():
        if arr is not None and _shape(arr) != tuple(expected_shape):
            raise ValueError(
                f"{name} has shape {_shape(arr)} but should have shape {tuple(expected_shape)} "
                "to match the Domain's interior points."
            )


class PDEProblem:
    """Elliptic problem L u = f on a uniform HPS Domain.

    L u = sum_k c_k(x) D_k u with D_k in [D_xx, D_xy, D_yy, D_x, D_y, I] (2D)
    or [D_xx, D_xy, D_yy, D_xz, D_yz, D_zz, D_x, D_y, D_z, I] (3D); D_xy is the
    single cross term D_x D_y.  ``use_ItI`` selects impedance-to-impedance
    merges with parameter ``eta`` (2D only).
    """

    def __init__(
        self,
        domain: Domain,
        source=None,
        D_xx_coefficients=None,
        D_xy_coefficients=None,
        D_xz_coefficients=None,
        D_yy_coefficients=None,
        D_yz_coefficients=None,
        D_zz_coefficients=None,
        D_x_coefficients=None,
        D_y_coefficients=None,
        D_z_coefficients=None,
        I_coefficients=None,
        use_ItI: bool = False,
        eta: float = None,
    ):
        self.domain = domain
        if isinstance(domain.root, DiscretizationNode3D):
            bool_2D = False
            if use_ItI:
                raise NotImplementedError("ItI merges are not supported for 3D problems.")
        else:
            bool_2D = True
            s = "z coefficients can not be set for 2D problems."
            for c in (D_xz_coefficients, D_yz_coefficients, D_zz_coefficients, D_z_coefficients):
                if c is not None:
                    raise ValueError(s)
        if use_ItI and eta is None:
            raise ValueError("eta must be specified when using ItI merges.")
        if use_ItI and not domain.bool_uniform:
            raise ValueError("ItI merges are only supported for uniform 2D problems.")
        if source is None and (not bool_2D or not domain.bool_uniform):
            raise ValueError("Source must be specified for non-uniform or 3D problems.")

        coeffs = dict(
            D_xx_coefficients=D_xx_coefficients,
            D_xy_coefficients=D_xy_coefficients,
            D_xz_coefficients=D_xz_coefficients,
            D_yy_coefficients=D_yy_coefficients,
            D_yz_coefficients=D_yz_coefficients,
            D_zz_coefficients=D_zz_coefficients,
            D_x_coefficients=D_x_coefficients,
            D_y_coefficients=D_y_coefficients,
            D_z_coefficients=D_z_coefficients,
            I_coefficients=I_coefficients,
        )
        check_input_shapes(source, use_ItI, domain.interior_points[..., 0].shape, **coeffs)
        for k, v in coeffs.items():
            setattr(self, k, v)
        self.source = source
        self.use_ItI = use_ItI
        self.eta = eta

        if domain.bool_uniform:
            half_side_len = (domain.root.xmax - domain.root.xmin) / (2 ** (domain.L + 1))
        else:
            # operators for half side 1; each leaf's size is applied to its data
            half_side_len = 1.0
            self.sidelens = np.array([lf.xmax - lf.xmin for lf in get_all_leaves(domain.root)])
        self.half_side_len = half_side_len
        p, q = domain.p, domain.q
        if bool_2D:
            self.D_x, self.D_y, self.D_xx, self.D_yy, self.D_xy = precompute_diff_operators_2D(p, half_side_len)
            if not use_ItI:
                self.P = precompute_P_2D_DtN(p, q)
                self.Q = precompute_Q_2D_DtN(p, q, self.D_x, self.D_y)
                if not domain.bool_uniform:
                    from ._adaptive import precompute_projection_ops_2D

                    self.L_2f1, self.L_1f2 = precompute_projection_ops_2D(q)
            else:
                self.P = precompute_P_2D_ItI(p, q)
                N_tilde = precompute_N_tilde_matrix_2D(self.D_x, self.D_y, p)
                self.G = precompute_G_2D_ItI(N_tilde, eta)
                N = precompute_N_matrix_2D(self.D_x, self.D_y, p)
                self.QH = precompute_QH_2D_ItI(N, p, q, eta)
        else:
            (
                self.D_x,
                self.D_y,
                self.D_z,
                self.D_xx,
                self.D_yy,
                self.D_zz,
                self.D_xy,
                self.D_xz,
                self.D_yz,
            ) = precompute_diff_operators_3D(p, half_side_len)
            self.P = precompute_P_3D_DtN(p, q)
            self.Q = precompute_Q_3D_DtN(p, q, self.D_x, self.D_y, self.D_z)
            if not domain.bool_uniform:
                from ._adaptive3d import precompute_projection_ops_3D

                self.L_4f1, self.L_1f4 = precompute_projection_ops_3D(q)

        self._dev_cache = {}
        self.reset()

    # -- solver state (``_pdeproblem.py:219-247``) ------------------------
    def reset(self) -> None:
        """Drop the stored solution operators."""
        self.Y = None
        self.v = None
        self.S_lst: List = []
        self.g_tilde_lst: List = []
        self.D_inv_lst: List = []
        self.BD_inv_lst: List = []
        self.Phi = None
        # factorisations kept on the device for the up pass, refinement and
        # the adjoint (LU of every leaf and merge matrix)
        self._factors = None
        self._precision = "fp64"
        self.refine_history: List = []
        # the refinement operators hold the coefficient fields
        for k in [k for k in self._dev_cache if k[0] == "refine"]:
            del self._dev_cache[k]

    def update_coefficients(
        self,
        source=None,
        D_xx_coefficients=None,
        D_xy_coefficients=None,
        D_xz_coefficients=None,
        D_yy_coefficients=None,
        D_yz_coefficients=None,
        D_zz_coefficients=None,
        D_x_coefficients=None,
        D_y_coefficients=None,
        D_z_coefficients=None,
        I_coefficients=None,
    ) -> None:
        """Replace the given coefficients and reset the solver (``_pdeproblem.py:249-333``)."""
        self.reset()
        new = dict(
            D_xx_coefficients=D_xx_coefficients,
            D_xy_coefficients=D_xy_coefficients,
            D_xz_coefficients=D_xz_coefficients,
            D_yy_coefficients=D_yy_coefficients,
            D_yz_coefficients=D_yz_coefficients,
            D_zz_coefficients=D_zz_coefficients,
            D_x_coefficients=D_x_coefficients,
            D_y_coefficients=D_y_coefficients,
            D_z_coefficients=D_z_coefficients,
            I_coefficients=I_coefficients,
        )
        check_input_shapes(self.source, self.use_ItI, self.domain.interior_points[..., 0].shape, **new)
        if source is not None:
            self.source = source
        for k, v in new.items():
            if v is not None:
                setattr(self, k, v)

    # -- helpers used by the stage functions ------------------------------
    def coefficient_list(self):
        """[(name, array)] of present coefficients in the reference's fixed order."""
        names = COEFF_NAMES_2D if self.domain.bool_2D else COEFF_NAMES_3D
        out = []
        for nm in names:
            arr = getattr(self, f"{nm}_coefficients")
            if arr is not None:
                out.append((nm, arr))
        return out

    def operator_names(self):
        return [n for n in (COEFF_NAMES_2D if self.domain.bool_2D else COEFF_NAMES_3D) if n != "I"]

    def device_ops(self, device, dtype):
        """Stacked diff operators and P/Q (or P/QH/G) on the device in `dtype`."""
        key = (str(device), dtype)
        if key not in self._dev_cache:
            names = self.operator_names()
            ops = np.stack([getattr(self, n) for n in names])

            def dev(x):
                return torch.as_tensor(np.ascontiguousarray(x)).to(device=device, dtype=dtype)

            d = {"diffops": dev(ops), "P": dev(self.P)}
            if self.use_ItI:
                d["Q"] = dev(self.QH)
                d["G"] = dev(self.G)
            else:
                d["Q"] = dev(self.Q)
            self._dev_cache[key] = d
        return self._dev_cache[key]
