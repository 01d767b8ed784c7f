"""``build_solver`` (``_build_solver.py:38-174``), uniform trees, device-resident.

Besides the reference's state (Y, v, S_lst, g_tilde_lst) the build keeps the
LU factors of every leaf and merge matrix plus each level's B (and D12, D21
for ItI) in ``pde_problem._factors``; they serve solve(..., source=...),
iterative refinement and the adjoint without refactorising.
"""

from __future__ import annotations

from ._discretization_tree import get_all_leaves

from ._lib import HPS_KIND_DTN, HPS_KIND_ITI, raise_on_info
from ._stages import cast
from .local_solve import _leaf
from .merge import _levels
from ._tables import quad_merge_tables


def _level_dtypes(n_levels, precision, hi, lo, fp64_top_levels):
    if precision == "fp64":
        return None
    return [hi if lvl >= n_levels - fp64_top_levels else lo for lvl in range(n_levels)]


def _compute_dtypes(pde_problem, precision, low_leaf=None):
    """(leaf factor dtype, leaf output storage dtype, FP64 dtype, low dtype).

    The mixed build factors the leaves in FP32 (complex64 for ItI) like the
    merges.  An ItI leaf matrix B = [G; A_int] is factored with its impedance
    rows scaled to the size of the collocation rows (_stages.iti_row_scale):
    unscaled, B's condition number at p = 16, L = 7 is ~2e6 and a complex64
    factorisation loses the k^2 (1 - q) term next to the O((p^2/h)^2)
    differentiation entries badly enough that refinement stops contracting;
    scaled, it is ~1e4 and the complex64 leaf maps are accurate to ~5e-6,
    which refinement corrects in two or three steps.  A 2D DtN leaf factors
    only its interior block A_ii (196 unknowns at p = 16), which costs little
    in FP64, while FP32 leaves leave the first solution accurate to only
    ~2e-3 and cost two more refinement steps (config B, 16 right-hand sides:
    34.3 ms with FP32 leaves, 29.4 ms with FP64 leaves); low_leaf=None
    therefore takes FP64 leaves for 2D DtN problems and low-precision leaves
    otherwise.  low_leaf=False / True forces FP64 / low-precision leaves."""
    from .local_solve import _compute_dtype
    from ._refine import LOW

    if precision not in ("fp64", "mixed"):
        raise ValueError(f"precision must be 'fp64' or 'mixed', got {precision!r}")
    hi = _compute_dtype(pde_problem, force_complex=pde_problem.use_ItI)
    if precision == "fp64":
        return hi, hi, hi, hi
    lo = LOW[hi]
    if low_leaf is None:
        low_leaf = pde_problem.use_ItI or not pde_problem.domain.bool_2D
    return (lo if low_leaf else hi), lo, hi, lo


def _release_state(pde_problem):
    """Drop a previous build's solver state before building again, so that a
    rebuild never holds two sets of factors (the reference overwrites the
    same attributes)."""
    pde_problem.Y = None
    pde_problem.v = None
    pde_problem.S_lst = []
    pde_problem.g_tilde_lst = []
    pde_problem._factors = None


def _merge_levels(pde_problem, T, h, return_top_T, dtypes=None):
    dom = pde_problem.domain
    kind = "iti" if pde_problem.use_ItI else "dtn"
    if dom.bool_2D:
        table_fn = lambda nec: quad_merge_tables(nec // 4, kind)  # noqa: E731
        nc = 4
    else:
        from ._tables import oct_merge_tables

        table_fn = lambda nec: oct_merge_tables(nec // 6)  # noqa: E731
        nc = 8
    levels, multi = _levels(T, h, dom.L, kind, nc, table_fn, keep_B=True, root_T=return_top_T, dtypes=dtypes)
    for r in levels:
        raise_on_info(r.info, "merge stage")
    return levels, multi


def build_solver(pde_problem, return_top_T: bool = False, compute_device=None, host_device=None,
                 precision: str = "fp64", fp64_top_levels: int = 0, low_precision_leaves=None):
    """Leaf solves and merges for the whole tree; stores Y, v, S_lst and
    g_tilde_lst in ``pde_problem``.  Returns the top-level T when
    ``return_top_T``.  ``compute_device`` may name the GPU (torch device);
    ``host_device`` is accepted for signature compatibility -- the solver
    state stays in HBM.

    precision="mixed" (uniform 2D and 3D trees) runs all but the top
    ``fp64_top_levels`` merge levels in FP32 (complex64 for ItI) and stores
    Y, v in that precision; the leaves are factored in FP32 / complex64
    (ItI and 3D; 2D DtN leaves stay in FP64, which saves refinement steps)
    unless ``low_precision_leaves`` says otherwise.  solve() then
    refines its solution to FP64 accuracy.  The default keeps no merge level
    in FP64: the refined error is set by the FP64 residual, and FP32 merges
    leave the first solution accurate enough (about 1e-6 at p = 16, L = 7,
    k = 100) for refinement to converge in one or two steps."""
    from . import profiling

    dom = pde_problem.domain
    _release_state(pde_problem)
    if not dom.bool_uniform:
        if precision != "fp64":
            raise NotImplementedError("adaptive trees are built in FP64")
        return _adaptive_build_solver(pde_problem, return_top_T, compute_device)
    leaf_dtype, store_dtype, hi, lo = _compute_dtypes(pde_problem, precision, low_precision_leaves)
    pde_problem._precision = precision
    dtypes = _level_dtypes(dom.L, precision, hi, lo, fp64_top_levels)
    if pde_problem.source is None:
        if not dom.bool_2D:
            raise ValueError(
                "Build stage for problems without source terms is only implemented for 2D uniform ItI problems."
            )
        return _nosource_build_solver(pde_problem, return_top_T, compute_device, leaf_dtype, dtypes, store_dtype)
    kind = HPS_KIND_ITI if pde_problem.use_ItI else HPS_KIND_DTN
    with profiling.stage("leaf"):
        res, v, h = _leaf(pde_problem, kind, compute_device, dtype=leaf_dtype)
    raise_on_info(res.info, "leaf stage")
    levels, multi = _merge_levels(pde_problem, res.T, h, return_top_T, dtypes)
    pde_problem.Y = cast(res.Y, store_dtype)
    pde_problem.v = cast(v, store_dtype)
    pde_problem.S_lst = [r.S for r in levels]
    pde_problem.g_tilde_lst = [r.gt if multi else r.gt[..., 0] for r in levels]
    if not dom.bool_2D:
        # the 3D root level is stored without its batch axis (merge/_uniform_3D_DtN.py:115-117)
        pde_problem.S_lst[-1] = pde_problem.S_lst[-1][0]
        pde_problem.g_tilde_lst[-1] = pde_problem.g_tilde_lst[-1][0]
    pde_problem._factors = {"leaf": res, "levels": levels}
    if return_top_T:
        return levels[-1].T[0]
    return None


def _adaptive_build_solver(pde_problem, return_top_T, compute_device):
    """Adaptive DtN build (``_build_solver.py:177-258``): leaf maps, then
    the merges, whose outputs live in the tree's nodes."""
    from ._adaptive import local_solve_stage_adaptive_2D_DtN, merge_stage_adaptive_2D_DtN
    from ._adaptive3d import merge_stage_adaptive_3D_DtN

    Y, T, v, h = local_solve_stage_adaptive_2D_DtN(pde_problem, device=compute_device)
    pde_problem.Y, pde_problem.v = Y, v
    for i, leaf in enumerate(get_all_leaves(pde_problem.domain.root)):
        leaf.data.Y, leaf.data.v = Y[i], v[i]
    if pde_problem.domain.bool_2D:
        merge_stage_adaptive_2D_DtN(pde_problem, T, h)
    else:
        merge_stage_adaptive_3D_DtN(pde_problem, T, h)
    pde_problem._precision = "fp64"
    if return_top_T:
        return pde_problem.domain.root.data.T
    return None


def _nosource_build_solver(pde_problem, return_top_T, compute_device, dtype=None, dtypes=None, store_dtype=None):
    """No-source build (``_build_solver.py:261-331``): leaf factors, S_lst and
    the merge factors; sources are supplied later to solve(..., source=...)."""
    from . import profiling
    from ._stages import leaf_solve
    from .local_solve import _compute_dtype, _resolve

    dom = pde_problem.domain
    kind = HPS_KIND_ITI if pde_problem.use_ItI else HPS_KIND_DTN
    device = _resolve(compute_device)
    if dtype is None:
        dtype = _compute_dtype(pde_problem, force_complex=(kind == HPS_KIND_ITI))
    with profiling.stage("leaf"):
        res = leaf_solve(pde_problem, kind, dtype, device, nsrc_zero=True)
    raise_on_info(res.info, "leaf stage")
    levels, _ = _merge_levels(pde_problem, res.T, None, return_top_T, dtypes)
    pde_problem.Y = cast(res.Y, store_dtype or res.Y.dtype)
    pde_problem.v = None
    pde_problem.S_lst = [r.S for r in levels]
    pde_problem.g_tilde_lst = []
    pde_problem._factors = {"leaf": res, "levels": levels}
    if return_top_T:
        return levels[-1].T[0]
    return None
