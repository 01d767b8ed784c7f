"""``Domain``: uniform quadtree/octree or adaptive quadtree discretization (``_domain.py:38-97``)."""

from __future__ import annotations

import numpy as np

from ._discretization_tree import DiscretizationNode2D, DiscretizationNode3D, get_all_leaves
from ._grid_creation import (
    compute_boundary_Gauss_points_uniform_2D,
    compute_boundary_Gauss_points_uniform_3D,
    compute_interior_Chebyshev_points_uniform_2D,
    compute_interior_Chebyshev_points_uniform_3D,
)


class Domain:
    """HPS discretization of a box.

    Args mirror the reference: p Chebyshev points per leaf side, q Gauss
    points per leaf side, the root node, and L levels of uniform refinement.
    Point arrays are float64 numpy arrays on the host:
    ``interior_points`` (n_leaves, p^d, d) and ``boundary_points``
    (n_bdry, d).  With L=None the tree is adaptive: the root's
    existing leaves are the leaves, and the side counts n_0..n_3 of every
    node are set.
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
            # adaptive tree: the leaves are the given tree's leaves
            from ._adaptive import (
                assign_side_counts,
                compute_boundary_Gauss_points_adaptive_2D,
                compute_interior_Chebyshev_points_adaptive_2D,
            )

            self.bool_uniform = False
            self.n_leaves = len(get_all_leaves(root))
            if self.bool_2D:
                assign_side_counts(root, q)
                self.interior_points = compute_interior_Chebyshev_points_adaptive_2D(root, p)
                self.boundary_points = compute_boundary_Gauss_points_adaptive_2D(root, q)
            else:
                from ._adaptive3d import (
                    assign_face_counts,
                    compute_boundary_Gauss_points_adaptive_3D,
                    compute_interior_Chebyshev_points_adaptive_3D,
                )

                assign_face_counts(root, q)
                self.interior_points = compute_interior_Chebyshev_points_adaptive_3D(root, p)
                self.boundary_points = compute_boundary_Gauss_points_adaptive_3D(root, q)
            return
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

    def get_adaptive_boundary_data_lst(self, f):
        """f evaluated on the root boundary points of each side S, E, N, W
        (``_domain.py:294-357``): a list of four arrays."""
        bp = self.boundary_points
        r = self.root
        if not self.bool_2D:
            return [f(bp[bp[:, 0] == r.xmin]), f(bp[bp[:, 0] == r.xmax]), f(bp[bp[:, 1] == r.ymin]),
                    f(bp[bp[:, 1] == r.ymax]), f(bp[bp[:, 2] == r.zmin]), f(bp[bp[:, 2] == r.zmax])]
        return [
            f(bp[bp[:, 1] == r.ymin]),
            f(bp[bp[:, 0] == r.xmax]),
            f(bp[bp[:, 1] == r.ymax]),
            f(bp[bp[:, 0] == r.xmin]),
        ]

    @classmethod
    def from_adaptive_discretization(cls, p: int, q: int, root, f, tol: float, use_level_restriction: bool = True,
                                     use_l_2_norm: bool = False) -> "Domain":
        """Refine root until each function in f is resolved to tol, then build
        the adaptive Domain (``_domain.py:359-434``)."""
        from ._adaptive import generate_adaptive_mesh_level_restriction_2D
        from ._adaptive3d import generate_adaptive_mesh_level_restriction

        gen = (generate_adaptive_mesh_level_restriction_2D if isinstance(root, DiscretizationNode2D)
               else generate_adaptive_mesh_level_restriction)
        for fn in (f if isinstance(f, list) else [f]):
            gen(root=root, f_fn=fn, tol=tol, p=p, q=q, restrict_bool=use_level_restriction, l2_norm=use_l_2_norm)
        return cls(p=p, q=q, root=root)

    def __repr__(self) -> str:
        return f"Domain(p={self.p}, q={self.q}, L={self.L}, {'2D' if self.bool_2D else '3D'}, n_leaves={self.n_leaves})"
