"""``solve`` (``_solve.py:18-93``), uniform trees."""

from __future__ import annotations

from .down_pass import down_pass_uniform_2D_DtN, down_pass_uniform_2D_ItI


def solve(pde_problem, boundary_data, source=None, compute_device=None, host_device=None):
    """Propagate root boundary data down the tree; returns u (n_leaves, p^d[, nsrc])
    as a device tensor."""
    dom = pde_problem.domain
    if source is not None:
        if not dom.bool_2D or not dom.bool_uniform:
            raise ValueError(
                "Source can only be specified for 2D uniform ItI problems. For other problems, "
                "the source must be specified at the time the solver is built."
            )
        raise NotImplementedError("solve(..., source=...) is not implemented yet")
    if not dom.bool_uniform:
        raise NotImplementedError("adaptive trees are outside this package's scope")
    if not dom.bool_2D:
        raise NotImplementedError("3D solve is not implemented yet")
    fn = down_pass_uniform_2D_ItI if pde_problem.use_ItI else down_pass_uniform_2D_DtN
    return fn(boundary_data, pde_problem.S_lst, pde_problem.g_tilde_lst, pde_problem.Y, pde_problem.v)
