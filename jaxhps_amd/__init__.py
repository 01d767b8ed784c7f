"""jaxhps_amd -- MI355X-native HPS direct solver hot path.

Drop-in for jaxhps's uniform-tree ``build_solver()`` / ``solve()`` with the
same ``Domain``, ``DiscretizationNode2D/3D`` and ``PDEProblem`` objects.  The
compute path is hand-written gfx950 HIP kernels in ``libhps_mi355x.so``
reached through its C ABI (``include/hps_mi355x.h``); torch provides device
memory and the stream.  There is no CPU fallback.
"""

from ._discretization_tree import DiscretizationNode2D, DiscretizationNode3D, get_all_leaves
from ._domain import Domain
from ._pdeproblem import PDEProblem
from ._build_solver import build_solver
from ._solve import solve
from ._lib import HPSLibraryError, HPSSingularError
from ._subtree_recomp import solve_subtree, upward_pass_subtree, downward_pass_subtree
from ._device_config import local_solve_chunksize_2D, local_solve_chunksize_3D

__all__ = [
    "Domain",
    "DiscretizationNode2D",
    "DiscretizationNode3D",
    "get_all_leaves",
    "PDEProblem",
    "build_solver",
    "solve",
    "solve_subtree",
    "local_solve_chunksize_2D",
    "local_solve_chunksize_3D",
    "upward_pass_subtree",
    "downward_pass_subtree",
    "HPSLibraryError",
    "HPSSingularError",
]
__version__ = "0.1"
