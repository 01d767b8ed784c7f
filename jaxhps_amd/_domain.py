"""``Domain``: uniform quadtree/octree discretization (``_domain.py:38-97``)."""

from __future__ import annotations

import numpy as np

from ._discretization_tree import DiscretizationNode2D, DiscretizationNode3D
from ._grid_creation import (
    compute_boundary_Gauss_points_uniform_2D,
    compute_boundary_Gauss_points_uniform_3D,
    compute_interior_Chebyshev_points_uniform_2D,
    compute_interior_Chebyshev_points_uniform_3D,
)


class Domain:
    """Uniform HPS discretization of a box.

    Args mirror the reference: p Chebyshev points per leaf side, q Gauss
    points per leaf side, the root node, and L levels of uniform refinement.
    Point arrays are float64 numpy arrays on the host:
    ``interior_points`` (n_leaves, p^d, d) and ``boundary_points``
    (n_bdry, d).  Adaptive trees (L=None) are not supported by this package.
    """

    def __init__(self, p: int, q: int, root, L: int | None = None):
        self.p = p
        self.q = q
        self.root = root
        self.L = L
        self.bool_2D = isinstance(root, DiscretizationNode2D)
        if not (self.bool_2D or isinstance(root, DiscretizationNode3D)):
            raise TypeError("root must be a DiscretizationNode2D or DiscretizationNode3D")
        if L is None:
            raise NotImplementedError(
                "adaptive discretizations are outside this package's scope; pass L for a uniform tree"
            )
        self.bool_uniform = True
        if self.bool_2D:
            self.interior_points = compute_interior_Chebyshev_points_uniform_2D(root, L, p)
            self.boundary_points = compute_boundary_Gauss_points_uniform_2D(root, L, q)
            self.n_leaves = 4**L
        else:
            self.interior_points = compute_interior_Chebyshev_points_uniform_3D(root, L, p)
            self.boundary_points = compute_boundary_Gauss_points_uniform_3D(root, L, q)
            self.n_leaves = 8**L

    def interp_to_interior_points(self, values, sample_points_x, sample_points_y, sample_points_z=None):
        """Values on meshgrid(x, y[, z], indexing="ij") -> (n_leaves, p^d) on
        the HPS grid (``_domain.py:99-195``); a torch tensor on the GPU."""
        from ._interp import interp_to_interior_points

        return interp_to_interior_points(self, values, sample_points_x, sample_points_y, sample_points_z)

    def interp_to_boundary_points(self, sample_points, sample_values, sample_f):
        """Not implemented in the reference either (``_domain.py:197-204``)."""
        raise NotImplementedError("interp_to_boundary_points is not implemented yet.")

    def interp_from_interior_points(self, samples, eval_points_x, eval_points_y, eval_points_z=None):
        """HPS samples (n_leaves, p^d[, nsrc]) -> (vals, target_pts) on a
        rectangular grid, in the reference's layouts (``_domain.py:206-294``)."""
        from ._interp import interp_from_interior_points

        return interp_from_interior_points(self, samples, eval_points_x, eval_points_y, eval_points_z)

    def __repr__(self) -> str:
        return f"Domain(p={self.p}, q={self.q}, L={self.L}, {'2D' if self.bool_2D else '3D'}, n_leaves={self.n_leaves})"
