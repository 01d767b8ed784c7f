"""``solve`` (``_solve.py:18-93``), uniform trees."""

from __future__ import annotations

from .down_pass import down_pass_uniform_2D_DtN, down_pass_uniform_2D_ItI, down_pass_uniform_3D_DtN
from .up_pass import up_pass_uniform_2D_DtN, up_pass_uniform_2D_ItI


def _up_then_down_pass(pde_problem, boundary_data, source):
    """Up pass for the new source, then the down pass (``_solve.py:115-151``)."""
    up = up_pass_uniform_2D_ItI if pde_problem.use_ItI else up_pass_uniform_2D_DtN
    down = down_pass_uniform_2D_ItI if pde_problem.use_ItI else down_pass_uniform_2D_DtN
    v, g_tilde_lst = up(source, pde_problem)
    return down(boundary_data, pde_problem.S_lst, g_tilde_lst, pde_problem.Y, v)


def _solve_mixed(pde_problem, boundary_data, source, refine_tol, max_refine):
    """First solve through the mixed-precision factors, then FP64 refinement."""
    from ._refine import refine_solve
    from .down_pass import _down

    if source is not None:
        up = up_pass_uniform_2D_ItI if pde_problem.use_ItI else up_pass_uniform_2D_DtN
        v, g_lst = up(source, pde_problem)
        f = source
    else:
        v, g_lst = pde_problem.v, pde_problem.g_tilde_lst
        f = pde_problem.source
    S_lst, table_fn = pde_problem.S_lst, None
    if not pde_problem.domain.bool_2D:
        from ._tables import oct_merge_tables

        # the 3D root level is stored without its batch axis
        S_lst = [*S_lst[:-1], S_lst[-1].unsqueeze(0)]
        g_lst = [*g_lst[:-1], g_lst[-1].unsqueeze(0)] if g_lst else g_lst
        table_fn = lambda ne: oct_merge_tables(ne // 24)  # noqa: E731
    u0, g_leaf = _down(boundary_data, S_lst, g_lst, pde_problem.Y, v, return_leaf_g=True, table_fn=table_fn)
    return refine_solve(pde_problem, boundary_data, u0, f, g_leaf0=g_leaf, tol=refine_tol, max_iter=max_refine)


def solve(pde_problem, boundary_data, source=None, compute_device=None, host_device=None,
          refine_tol: float = 1e-11, max_refine: int = 10):
    """Propagate root boundary data down the tree; returns u (n_leaves, p^d[, nsrc])
    as a device tensor.

    After build_solver(..., precision="mixed") the solution is refined in FP64
    until the relative update is below ``refine_tol`` (at most ``max_refine``
    iterations; the history is left in ``pde_problem.refine_history``)."""
    dom = pde_problem.domain
    if getattr(pde_problem, "_precision", "fp64") == "mixed":
        if source is not None and (not dom.bool_2D or not dom.bool_uniform):
            raise ValueError(
                "Source can only be specified for 2D uniform ItI problems. For other problems, "
                "the source must be specified at the time the solver is built."
            )
        return _solve_mixed(pde_problem, boundary_data, source, refine_tol, max_refine)
    if source is not None:
        if not dom.bool_2D or not dom.bool_uniform:
            raise ValueError(
                "Source can only be specified for 2D uniform ItI problems. For other problems, "
                "the source must be specified at the time the solver is built."
            )
        return _up_then_down_pass(pde_problem, boundary_data, source)
    if not dom.bool_uniform:
        from ._adaptive import down_pass_adaptive_2D_DtN

        return down_pass_adaptive_2D_DtN(pde_problem, boundary_data)
    if not dom.bool_2D:
        fn = down_pass_uniform_3D_DtN
    else:
        fn = down_pass_uniform_2D_ItI if pde_problem.use_ItI else down_pass_uniform_2D_DtN
    return fn(boundary_data, pde_problem.S_lst, pde_problem.g_tilde_lst, pde_problem.Y, pde_problem.v)
