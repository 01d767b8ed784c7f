#!/usr/bin/env python
"""Benchmark: HPS build_solver + solve throughput (DOF/s) on one MI355X.

A step is one build_solver() + solve() of the workload (all right-hand
sides), with coefficients, source and boundary data already in HBM.  DOF =
n_leaves * p^d.  Prints ONE JSON line (rank 0).

Two paths are timed on the same inputs: the mixed-precision build
(complex64 / FP32 merge factors at every level by default, leaves as the
library chooses) whose solve is refined to FP64 accuracy -- the reported
value -- and the plain FP64 path.  The line carries the relative error of
the mixed result against the FP64 one, the speed-up, the same mixed build
with its top two merge levels in FP64 (variant_fp64_top2), and, with the CPU
baseline, both paths' errors against the CPU oracle and the exact discrete
solution at L = 4 of the same workload (oracle_check).  The process exits
with status 3 after printing the line when the mixed result misses 1e-8
against the FP64 path or the oracle, the FP64 path misses 1e-10 against the
oracle, or the refinement did not converge.

Default workload: 2D Helmholtz, ItI merges, k = 100, p = 16, L = 7 (16384
leaves, 4,194,304 DOF).  --config A, B or D selects the 2D Poisson DtN
workload (p = 8, L = 3), the 2D variable-coefficient DtN workload (p = 16,
L = 6, 16 RHS) or the 3D variable-coefficient DtN workload (p = 10, L = 3,
4 RHS).  The default line also carries short timed runs of A, B and D
(other_configs) and the adjoint's gradients per second.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="C", choices=["A", "B", "C", "D"])
    ap.add_argument("--L", type=int, default=None, help="override tree depth (testing only)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--profile-only", action="store_true", help="skip the CPU baseline and extra passes")
    ap.add_argument("--precision", default="both", choices=["both", "mixed", "fp64"],
                    help="which path(s) to time; the value is the mixed path's unless only fp64 runs")
    ap.add_argument("--fp64-top-levels", type=int, default=0,
                    help="merge levels at the top kept in FP64 by the mixed path (the library's default, 0: "
                         "complex64 merges suffice for <= 1e-8, the refined error is set by the FP64 residual)")
    ap.add_argument("--leaf-precision", default="auto", choices=["auto", "low", "high"],
                    help="mixed path leaf factors: the library's default (auto), FP32/complex64 (low) or FP64 (high)")
    ap.add_argument("--no-oracle-check", action="store_true",
                    help="skip the accuracy check of both GPU paths against the CPU oracle at L = 4")
    ap.add_argument("--no-variants", action="store_true",
                    help="skip timing the mixed path with the top two merge levels in FP64")
    ap.add_argument("--refine-tol", type=float, default=1e-10,
                    help="refinement stops once the (measured or contraction-estimated) relative update is below "
                         "this; 1e-10 leaves a 100x margin to the 1e-8 bar, and the line reports the measured error")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint grads/s measurement")
    ap.add_argument("--no-other-configs", action="store_true",
                    help="skip the short timed runs of configs A, B and D on the default line")
    ap.add_argument("--no-surrogate", action="store_true",
                    help="skip the surrogate-in-the-gradient-loop measurement of the adjoint line")
    ap.add_argument("--adjoint-L", type=int, default=6, help="tree depth of the adjoint workload (testing only)")
    ap.add_argument("--workload", default="hps",
                    choices=["hps", "scattering", "interp", "adaptive", "adaptive3d", "subtree"],
                    help="hps: the headline build+solve; the others time the exterior coupling, the grid "
                         "interpolation and the adaptive 2D build+solve")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain device copy of known size (hps_convert, complex128 -> complex128, "
                         "256 MiB, 10 times): the reference for scaling HBM byte counters")
    return ap.parse_args()


def make_workload(cfg, L=None):
    from jaxhps_amd import workloads as wl

    if cfg == "A":
        dom, kw, g = wl.config_A(L=L or 3)
        desc = dict(workload="2D Poisson DtN p=8 (manufactured sin sin)", L=dom.L, n_leaves=dom.n_leaves, p=8, q=6,
                    nsrc=1)
    elif cfg == "C":
        dom, kw, g = wl.config_C(L=L or 7)
        desc = dict(workload="2D Helmholtz ItI k=100 p=16", L=dom.L, n_leaves=dom.n_leaves, p=16, q=14, nsrc=1)
    elif cfg == "B":
        dom, kw, g = wl.config_B(L=L or 6)
        desc = dict(workload="2D var-coeff elliptic DtN p=16", L=dom.L, n_leaves=dom.n_leaves, p=16, q=14, nsrc=16)
    else:
        dom, kw, g = wl.config_D(L=L or 3)
        desc = dict(workload="3D var-coeff Poisson DtN p=10", L=dom.L, n_leaves=dom.n_leaves, p=10, q=8, nsrc=4)
    return dom, kw, g, desc


def other_configs_bench(refine_tol):
    """Short timed runs of the other workloads for the default line: config A
    (FP64, the plain method), B and D (mixed with FP64 refinement, and FP64),
    each at its full size, inputs resident in HBM, one warmup step, then a
    few timed build+solve steps; the mixed result's error against the FP64
    one and the refinement's iteration count come with each."""
    import numpy as np
    import torch

    import jaxhps_amd as jh

    dev = torch.device("cuda", 0)
    out = {}
    for cfg, steps in (("A", 5), ("B", 3), ("D", 2)):
        dom, kw, g, desc = make_workload(cfg)
        dkw = {k: (torch.as_tensor(v).to(dev) if isinstance(v, np.ndarray) else v) for k, v in kw.items()}
        g_dev = torch.as_tensor(g).to(dev)
        dof = dom.n_leaves * dom.p ** (2 if dom.bool_2D else 3)
        rec = dict(desc, dof=dof, steps=steps, warmup=1)
        us = {}
        for prec in (["fp64"] if cfg == "A" else ["mixed", "fp64"]):
            pde = jh.PDEProblem(dom, **dkw)
            bkw = {} if prec == "fp64" else dict(precision="mixed", fp64_top_levels=0)

            def step():
                jh.build_solver(pde, **bkw)
                if prec == "mixed":
                    return jh.solve(pde, g_dev, refine_tol=refine_tol)
                return jh.solve(pde, g_dev)

            step()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(steps):
                u = step()
            torch.cuda.synchronize()
            ms = (time.perf_counter() - t0) / steps * 1e3
            rec[prec] = {"ms_per_step": ms, "value": dof / (ms / 1e3), "unit": "DOF/s"}
            if prec == "mixed":
                rec[prec]["refine_iterations"] = len(pde.refine_history)
                rec[prec]["converged"] = bool(pde.refine_converged)
            us[prec] = u
            del pde
        if "mixed" in us:
            rec["speedup_vs_fp64"] = rec["fp64"]["ms_per_step"] / rec["mixed"]["ms_per_step"]
            rec["rel_err_vs_fp64"] = float(torch.linalg.vector_norm(us["mixed"] - us["fp64"])
                                           / torch.linalg.vector_norm(us["fp64"]))
        out[cfg] = rec
        del us
        torch.cuda.empty_cache()
    return out


def _oracle_setup_2d(dom, kw):
    import numpy as np

    from oracle import hps2d
    from oracle import spectral as osp

    names = {"D_xx_coefficients": "D_xx", "D_yy_coefficients": "D_yy", "D_x_coefficients": "D_x",
             "D_y_coefficients": "D_y", "I_coefficients": "I"}
    iti = kw.get("use_ItI", False)
    ops = osp.leaf_ops_2D(dom.p, dom.q, 2.0 / 2 ** (dom.L + 1), eta=kw.get("eta") if iti else None)
    coeffs = {names[k]: np.asarray(v) for k, v in kw.items() if k in names}
    src = np.asarray(kw["source"])
    src = src if src.ndim == 3 else src[..., None]
    return ops, coeffs, (src.astype(complex) if iti else src), iti, hps2d


def cpu_baseline(cfg):
    """The FP64 oracle (numpy/LAPACK on the host cores) on the workload.

    Config B (L = 6) runs whole.  Config C (L = 7, 16384 leaves) is timed by
    parts at full size and the parts are summed: the leaf stage on 512 of
    the leaves (x 32); at every merge level one 4:1 merge of the level's real
    size (children's maps are synthetic at levels >= 2: a merge's cost does
    not depend on the values) x the level's merge count; the down pass the
    same way; the leaf apply on the sampled leaves (x 16).  Config D (3D)
    runs whole at L = 2."""
    import numpy as np

    from jaxhps_amd import workloads as wl

    if cfg == "D":
        return _cpu_baseline_3d()
    cores = len(os.sched_getaffinity(0))
    if cfg == "A":
        dom, kw, g = wl.config_A()
        ops, coeffs, src, _, hps2d = _oracle_setup_2d(dom, kw)

        def run_a():
            A = hps2d.assemble_diff_operator(coeffs, ops)
            Y, T, v, h = hps2d.leaf_dtn(A, src, ops["P"], ops["Q"])
            S, gl, _, _ = hps2d.merge_stage(T, h, dom.L, "dtn")
            return hps2d.down_pass(g[:, None], S, gl, Y, v, "dtn")[0]

        t, n = _time_bounded(run_a, 10.0)
        dof = dom.n_leaves * dom.p**2
        return {"value": dof / t, "unit": "DOF/s", "cores": cores, "kind": "port",
                "sample": f"FP64 numpy oracle, the whole build+solve of the workload (L=3, 64 leaves, {dof} DOF), "
                          f"median of {n} runs, BLAS threads <= {cores}"}
    if cfg == "B":
        dom, kw, g = wl.config_B(L=6)
        ops, coeffs, src, _, hps2d = _oracle_setup_2d(dom, kw)

        def run():
            A = hps2d.assemble_diff_operator(coeffs, ops)
            Y, T, v, h = hps2d.leaf_dtn(A, src, ops["P"], ops["Q"])
            S, gl, _, _ = hps2d.merge_stage(T, h, dom.L, "dtn")
            return hps2d.down_pass(g, S, gl, Y, v, "dtn")[0]

        t0 = time.perf_counter()
        run()
        t = time.perf_counter() - t0
        dof = dom.n_leaves * dom.p**2
        return {"value": dof / t, "unit": "DOF/s", "cores": cores, "kind": "port",
                "sample": f"FP64 numpy oracle, the whole build+solve of the workload (L=6, {dom.n_leaves} leaves, "
                          f"{dof} DOF, 16 RHS), one run, BLAS threads <= {cores}"}
    dom, kw, g = wl.config_C(L=7)
    ops, coeffs, src, _, hps2d = _oracle_setup_2d(dom, kw)
    nl, L, nq = dom.n_leaves, dom.L, 4 * dom.q
    ns = 512
    parts = {}
    t0 = time.perf_counter()
    A = hps2d.assemble_diff_operator({k: v[:ns] for k, v in coeffs.items()}, ops)
    T, Y, h, v = hps2d.leaf_iti(A, src[:ns], ops["P_iti"], ops["QH"], ops["G"])
    parts["leaf"] = (time.perf_counter() - t0) * nl / ns
    rng = np.random.default_rng(0)
    down = 0.0
    for lvl in range(1, L + 1):
        ne = nq * 2 ** (lvl - 1)
        if lvl == 1:
            kids = [T[i] for i in range(4)]
            hs = [h[i] for i in range(4)]
        else:
            kids = [(rng.standard_normal((ne, ne)) + 1j * rng.standard_normal((ne, ne))) * (0.5 / np.sqrt(ne))
                    for _ in range(4)]
            hs = [rng.standard_normal((ne, 1)) + 0j for _ in range(4)]
        reps = max(1, min(64, 2 ** (2 * (L - lvl)) // 64))
        t0 = time.perf_counter()
        for _ in range(reps):
            S, Tn, hn, gt = hps2d.quad_merge_iti(*kids, *hs)
        parts[f"merge_level_{lvl}"] = (time.perf_counter() - t0) / reps * nl / 4**lvl
        gpar = rng.standard_normal((2 * ne, 1)) + 0j
        t0 = time.perf_counter()
        for _ in range(reps):
            hps2d.propagate_down_iti(S, gpar, gt)
        down += (time.perf_counter() - t0) / reps * nl / 4**lvl
        del kids, S, Tn
    gl = rng.standard_normal((ns, nq, 1)) + 0j
    t0 = time.perf_counter()
    _ = Y @ gl + v
    parts["down_pass"] = down + (time.perf_counter() - t0) * nl / ns
    total = sum(parts.values())
    dof = nl * dom.p**2
    return {"value": dof / total, "unit": "DOF/s", "cores": cores, "kind": "port",
            "seconds_extrapolated": total, "parts_s": {k: round(v, 3) for k, v in parts.items()},
            "sample": f"FP64 numpy oracle timed by parts at the workload's full size (L=7, {nl} leaves, {dof} DOF): "
                      f"leaf stage on {ns} leaves x {nl // ns}; one 4:1 merge and one down-pass step per level at "
                      "its real size x the level's merge count; BLAS threads <= " + str(cores)}


def oracle_check(cfg, mixed_kwargs, refine_tol, L=4):
    """Both GPU paths on the workload's own generator at L = 4 against the
    FP64 CPU oracle (stage by stage) and against the exact discrete solution
    of the same global system (sparse LU + extended-precision refinement).
    Relative 2-norm errors."""
    import numpy as np
    import torch

    import jaxhps_amd as jh
    from jaxhps_amd import workloads as wl

    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import config_oracle as co

    dom, kw, g = {"A": wl.config_A, "B": wl.config_B, "C": wl.config_C}[cfg](L=min(L, 3) if cfg == "A" else L)
    if cfg == "B":
        kw = dict(kw)
        kw["source"] = np.random.default_rng(5).standard_normal(kw["source"].shape)
    t0 = time.perf_counter()
    u_stage = co.stage_solution(dom, kw, g)
    u_exact = co.exact_solution(dom, kw, g)
    t_cpu = time.perf_counter() - t0
    out = {"L": L, "n_leaves": dom.n_leaves, "oracle_seconds": t_cpu}
    for name, bkw in (("fp64", {}), ("mixed", mixed_kwargs)):
        pde = jh.PDEProblem(dom, **kw)
        jh.build_solver(pde, **bkw)
        u = jh.solve(pde, g, refine_tol=refine_tol) if bkw else jh.solve(pde, g)
        u = u.cpu().numpy().reshape(u_stage.shape)
        out[f"{name}_vs_oracle"] = co.rel2(u, u_stage)
        out[f"{name}_vs_exact"] = co.rel2(u, u_exact)
    out["oracle_vs_exact"] = co.rel2(u_stage, u_exact)
    return out


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the rocprofv3 PMC summaries under
    profiles/ (a 'traffic: <kernel> <bytes>' line), or (None, None)."""
    import glob

    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "*.md"))):
        with open(path) as fh:
            for line in fh:
                parts = line.split()
                if len(parts) == 3 and parts[0] == "traffic:" and parts[1] == kernel:
                    return float(parts[2]), os.path.relpath(path, ROOT)
    return None, None


def _time_bounded(run, budget_s=20.0):
    run()
    times = []
    t_end = time.time() + budget_s
    while len(times) < 3 and (time.time() < t_end or not times):
        t0 = time.perf_counter()
        run()
        times.append(time.perf_counter() - t0)
    return sorted(times)[len(times) // 2], len(times)


def _cpu_baseline_3d():
    from jaxhps_amd import workloads as wl
    from oracle import hps2d, hps3d

    L = 2
    dom, kw, g = wl.config_D(L=L)
    p, q = dom.p, dom.q
    ops = hps3d.leaf_ops_3D(p, q, 2.0 / 2 ** (L + 1))
    names = {"D_xx_coefficients": "D_xx", "D_yy_coefficients": "D_yy", "D_zz_coefficients": "D_zz",
             "D_x_coefficients": "D_x", "D_y_coefficients": "D_y", "D_z_coefficients": "D_z"}
    coeffs = {names[k]: v for k, v in kw.items() if k in names}
    order = ("D_xx", "D_xy", "D_yy", "D_xz", "D_yz", "D_zz", "D_x", "D_y", "D_z", "I")

    def run():
        A = hps2d.assemble_diff_operator(coeffs, ops, order=order)
        Y, T, v, h = hps2d.leaf_dtn(A, kw["source"], ops["P"], ops["Q"])
        S, gl, _, _ = hps3d.merge_stage_3d(T, h, L)
        return hps3d.down_pass_3d(g, S, gl, Y, v)[0]

    t, n = _time_bounded(run, 30.0)
    dof = dom.n_leaves * p**3
    cores = len(os.sched_getaffinity(0))
    return {"value": dof / t, "unit": "DOF/s", "cores": cores, "kind": "port",
            "sample": f"FP64 numpy oracle, full 3D build+solve at L={L} ({dom.n_leaves} leaves, {dof} DOF, 4 RHS), "
                      f"median of {n} runs, BLAS threads <= {cores}"}


def _sine_field():
    """q = theta . basis with the transposed product written as a weighted row
    sum (the library GEMV of torch on this image takes ~170 ms for the
    25 x 1M transposed product)."""
    import torch

    class SineField(torch.autograd.Function):
        @staticmethod
        def forward(ctx, theta, basis_flat):
            ctx.save_for_backward(basis_flat)
            return theta @ basis_flat

        @staticmethod
        def backward(ctx, g):
            (basis_flat,) = ctx.saved_tensors
            return (basis_flat * g).sum(dim=1), None

    return SineField.apply


def adjoint_bench(L=6, warmup=1, reps=2):
    """Inverse-scattering gradients: config E's problem (k = 40, p = 16, 25 sine
    coefficients, 8 seeded thetas); one gradient = build + solve + adjoint of
    J(theta) = ||u(q_theta) - u_obs||^2 for one theta.  Returns grads/s over the
    batch of 8, timed after warmup."""
    import numpy as np
    import torch

    import jaxhps_amd as jh
    from jaxhps_amd import workloads as wl
    from jaxhps_amd.adjoint import differentiable_solve

    k = 40.0
    dom, kw, g, basis, thetas, theta_star = wl.config_E(L=L, p=16, k=k)
    dev = torch.device("cuda", 0)
    basis_t = torch.as_tensor(basis, device=dev)
    basis_flat = basis_t.reshape(basis_t.shape[0], -1).contiguous()
    field_shape = basis_t.shape[1:]
    sine_field = _sine_field()
    g_dev = torch.as_tensor(g, device=dev)
    dkw = {key: (torch.as_tensor(v, device=dev) if isinstance(v, np.ndarray) else v) for key, v in kw.items()}
    ones = torch.ones(dom.interior_points.shape[:2], dtype=torch.float64, device=dev)
    pde = jh.PDEProblem(dom, I_coefficients=ones, **dkw)
    with torch.no_grad():
        I_star = k**2 * (1.0 - (torch.as_tensor(theta_star, device=dev) @ basis_flat).reshape(field_shape))
        u_obs = differentiable_solve(pde, g_dev, I=I_star).clone()
    thetas_t = torch.as_tensor(thetas, device=dev)

    def batch(ckpt):
        grads = []
        for th in thetas_t:
            th = th.clone().requires_grad_(True)
            I = k**2 * (1.0 - sine_field(th, basis_flat).reshape(field_shape))
            u = differentiable_solve(pde, g_dev, checkpoint=ckpt, I=I)
            J = ((u - u_obs).abs() ** 2).sum()
            J.backward()
            grads.append(th.grad)
        return torch.stack(grads)

    def held_between_forward_and_backward(ckpt):
        """Device bytes the forward pass leaves allocated for the backward."""
        pde.reset()
        th = thetas_t[0].clone().requires_grad_(True)
        torch.cuda.synchronize()
        m0 = torch.cuda.memory_allocated(dev)
        u = differentiable_solve(pde, g_dev, checkpoint=ckpt,
                                 I=k**2 * (1.0 - sine_field(th, basis_flat).reshape(field_shape)))
        J = ((u - u_obs).abs() ** 2).sum()
        torch.cuda.synchronize()
        held = torch.cuda.memory_allocated(dev) - m0
        J.backward()
        return int(held)

    res = {}
    for ckpt in (False, True):
        for _ in range(warmup):
            batch(ckpt)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            grads = batch(ckpt)
        torch.cuda.synchronize()
        t = (time.perf_counter() - t0) / reps
        res[ckpt] = (t, grads, held_between_forward_and_backward(ckpt))
    t, grads, mem = res[False]
    tc, gc, memc = res[True]
    return {"metric": "adjoint grads/s", "value": thetas.shape[0] / t, "unit": "grads/s",
            "ms_per_grad": t / thetas.shape[0] * 1e3, "batch": int(thetas.shape[0]),
            "config": {"workload": "2D Helmholtz inverse scattering ItI k=40 p=16, 25 sine coefficients",
                       "L": L, "n_leaves": dom.n_leaves},
            "dtype": "c128", "grad_norm": float(grads.norm()), "held_device_bytes": mem,
            "checkpointed": {"value": thetas.shape[0] / tc, "ms_per_grad": tc / thetas.shape[0] * 1e3,
                             "held_device_bytes": memc,
                             "max_rel_diff_vs_stored": float((gc - grads).abs().max() / grads.abs().max())}}


def surrogate_loop_bench(L=6, iters=4, trials=3, n_fields=4, train_steps=1000):
    """The learned leaf-ItI surrogate inside config E's gradient loop.

    Gradient descent on J(theta) = ||u(q_theta) - u_obs||^2 with a line
    search: each iteration takes one gradient through the plain adjoint,
    then evaluates J at `trials` step lengths and takes the best, evaluating
    J either with the plain FP64 build + solve (surrogate off)
    or with the surrogate forward solve (surrogate on: MLP leaf maps, each
    checked on a random boundary probe with fallback to the plain leaf solve,
    complex64 merges, plain interior leaf solves with the shared q = 0
    factors).  The plain method judges the result: J at the final theta of
    both loops is evaluated with the plain solve.  Also: objective
    evaluations per second each way over the 8 seeded thetas and the largest
    relative difference of J."""
    import numpy as np
    import torch

    import jaxhps_amd as jh
    from jaxhps_amd import workloads as wl
    from jaxhps_amd.adjoint import differentiable_solve
    from jaxhps_amd.surrogate import LeafItISurrogate, SharedLeafSolver, leaf_training_data, surrogate_solve

    k = 40.0
    dev = torch.device("cuda", 0)
    dom, kw, g, basis, thetas, theta_star = wl.config_E(L=L, p=16, k=k)
    basis_flat = torch.as_tensor(basis, device=dev).reshape(basis.shape[0], -1).contiguous()
    shape = basis.shape[1:]
    g_dev = torch.as_tensor(g, device=dev)
    dkw = {key: (torch.as_tensor(v, device=dev) if isinstance(v, np.ndarray) else v) for key, v in kw.items()}
    pde = jh.PDEProblem(dom, I_coefficients=torch.ones(shape, dtype=torch.float64, device=dev), **dkw)

    def I_of(th):
        return k**2 * (1.0 - (th @ basis_flat).reshape(shape))

    t0 = time.perf_counter()
    q, T, T0 = leaf_training_data(n_fields=n_fields, L=L, p=16, k=k, seed=1, device=dev)
    model = LeafItISurrogate(n_in=256, width=512, nq=56, rank=8, seed=0, device=dev)
    model.fit(q, T, T0, steps=train_steps, batch=256, lr=1e-3)
    solver = SharedLeafSolver(pde, I_ref=k**2, device=dev)
    torch.cuda.synchronize()
    train_s = time.perf_counter() - t0
    stats = {"passed": [], "fallback_leaves": 0}

    with torch.no_grad():
        u_obs = differentiable_solve(pde, g_dev, I=I_of(torch.as_tensor(theta_star, device=dev))).clone()

    def J_plain(th):
        with torch.no_grad():
            u = differentiable_solve(pde, g_dev, I=I_of(th))
            return float(((u - u_obs).abs() ** 2).sum())

    def J_sur(th):
        with torch.no_grad():
            r = surrogate_solve(model, solver, pde, g_dev, I_of(th), k)
            stats["passed"].append(float(r.passed.double().mean()))
            stats["fallback_leaves"] += int((~r.passed).sum())
            return float(((r.u - u_obs).abs() ** 2).sum())

    def grad(th):
        th = th.clone().requires_grad_(True)
        u = differentiable_solve(pde, g_dev, I=I_of(th))
        J = ((u - u_obs).abs() ** 2).sum()
        J.backward()
        return float(J.detach()), th.grad.detach()

    thetas_t = torch.as_tensor(thetas, device=dev)

    def loop(evalfn):
        # each iteration: one gradient, then `trials` step lengths evaluated
        # and the best one taken
        th = thetas_t[0].clone()
        n_eval = 0
        for _ in range(iters):
            J0, gr = grad(th)
            alpha = 0.1 / float(gr.norm())
            cands = [th - alpha * 0.5**t * gr for t in range(trials)]
            Js = [evalfn(c) for c in cands]
            n_eval += trials
            th = cands[int(np.argmin(Js))]
        return th, n_eval

    out = {"training_seconds": train_s, "train_samples": int(q.shape[0]), "iters": iters, "trials": trials}
    for name, fn in (("plain", J_plain), ("surrogate", J_sur)):
        fn(thetas_t[1])  # warm
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        Js = [fn(th) for th in thetas_t]
        torch.cuda.synchronize()
        out[f"{name}_evals_per_s"] = len(Js) / (time.perf_counter() - t0)
        out[f"J_{name}"] = Js
        loop(fn)  # warm
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        th_end, n_eval = loop(fn)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        out[f"{name}_loop"] = {"iterations_per_s": iters / dt, "seconds": dt, "objective_evals": n_eval,
                               "final_J_plain": J_plain(th_end)}
    Jp, Js = np.array(out.pop("J_plain")), np.array(out.pop("J_surrogate"))
    out["max_rel_J_diff"] = float(np.abs(Js - Jp).max() / np.abs(Jp).max())
    out["mean_surrogate_leaf_fraction"] = float(np.mean(stats["passed"]))
    out["fallback_leaves_total"] = stats["fallback_leaves"]
    out["eval_speedup"] = out["surrogate_evals_per_s"] / out["plain_evals_per_s"]
    return out


def _line(metric, unit, value, ms, steps, warmup, dtype, config, roofline, kernels, cpu):
    return {"metric": metric, "value": value, "unit": unit, "n_gpus": 1, "steps": steps, "warmup": warmup,
            "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": dtype,
            "data": "synthetic", "config": config, "roofline": roofline, "kernels": kernels, "cpu_baseline": cpu}


def _timed_steps(step, steps, warmup):
    import torch

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps * 1e3


def bench_scattering(args, n=7168, nsrc=16):
    """Exterior coupling of a top-level operator: ItI -> DtN and the
    boundary-integral solve for nsrc incident waves (n = 7168 is the root
    boundary of the 2D p = 16, L = 7 tree).  Step = get_DtN_from_ItI +
    get_scattering_uscat_impedance on resident inputs."""
    import numpy as np
    import torch

    from jaxhps_amd import profiling
    from jaxhps_amd.scattering import get_DtN_from_ItI, get_scattering_uscat_impedance

    dev = torch.device("cuda", 0)
    g = torch.Generator(device=dev).manual_seed(0)

    def crand(scale):
        return torch.complex(torch.randn(n, n, generator=g, device=dev, dtype=torch.float64),
                             torch.randn(n, n, generator=g, device=dev, dtype=torch.float64)) * scale

    R, S, D = crand(0.5 / np.sqrt(n)), crand(1.0 / n), crand(1.0 / n)
    bdry = torch.as_tensor(np.random.default_rng(1).uniform(-1, 1, (n, 2))).to(dev)
    dirs = torch.linspace(0, 2 * np.pi, nsrc + 1, dtype=torch.float64, device=dev)[:-1]
    k = 100.0

    def step():
        T = get_DtN_from_ItI(R, k)
        return get_scattering_uscat_impedance(S, D, T, dirs, bdry, k, k)

    ms = _timed_steps(step, args.steps, args.warmup)
    _, kernels, roof = profiling.instrumented(step)
    cpu = None
    if not args.no_cpu_baseline:
        from oracle import scattering as osc

        m = 1792
        rng = np.random.default_rng(2)
        Rh = (rng.standard_normal((m, m)) + 1j * rng.standard_normal((m, m))) * 0.5 / np.sqrt(m)
        Sh = (rng.standard_normal((m, m)) + 1j * rng.standard_normal((m, m))) / m
        Dh = (rng.standard_normal((m, m)) + 1j * rng.standard_normal((m, m))) / m
        bh = rng.uniform(-1, 1, (m, 2))
        dh = np.linspace(0, 2 * np.pi, nsrc + 1)[:-1]
        t, reps = _time_bounded(lambda: osc.uscat_impedance(Sh, Dh, osc.dtn_from_iti(Rh, k), dh, bh, k, k), 20.0)
        cores = len(os.sched_getaffinity(0))
        cpu = {"value": (1.0 / t) * (m / n) ** 3, "unit": "couplings/s", "cores": cores, "kind": "port",
               "sample": f"numpy oracle at n = {m} (median of {reps}), scaled by (n/{m})^3 to n = {n}"}
    return _line(f"exterior coupling (ItI->DtN + BIE solve) per second at n_bdry={n}", "couplings/s",
                 1e3 / ms, ms, args.steps, args.warmup, "c128",
                 {"workload": "exterior scattering coupling", "n_bdry": n, "nsrc": nsrc}, roof, kernels, cpu)


def bench_interp(args, ngrid=1024, nsamp=64):
    """Both interpolation directions on the 2D p = 16, L = 7 HPS grid (16384
    leaves, 4.2M points): from nsamp x nsamp samples at Chebyshev abscissae
    (one global interpolant per axis, as the reference builds it) to the HPS
    grid, and from the HPS grid to an ngrid x ngrid uniform grid.  Step =
    interp_to_interior_points + interp_from_interior_points."""
    import numpy as np
    import torch

    import jaxhps_amd as jh
    from jaxhps_amd import profiling

    dev = torch.device("cuda", 0)
    dom = jh.Domain(p=16, q=14, root=jh.DiscretizationNode2D(-1.0, 1.0, -1.0, 1.0), L=7)
    xs = np.linspace(-1, 1, ngrid)
    cs = -np.cos(np.pi * np.arange(nsamp) / (nsamp - 1))
    vals = torch.as_tensor(np.sin(3 * cs)[:, None] * np.cos(2 * cs)[None, :]).to(dev)
    samples = torch.as_tensor(np.sin(dom.interior_points[..., 0]) * dom.interior_points[..., 1]).to(dev)

    def step():
        a = dom.interp_to_interior_points(vals, cs, cs)
        b, _ = dom.interp_from_interior_points(samples, xs, xs)
        return a, b

    ms = _timed_steps(step, args.steps, args.warmup)
    _, kernels, roof = profiling.instrumented(step)
    npts = dom.n_leaves * 256 + ngrid * ngrid
    cpu = None
    if not args.no_cpu_baseline:
        from oracle import interp as oi

        m = 64
        t, reps = _time_bounded(lambda: oi.interp_from_interior_2D((-1, 1, -1, 1), 7, 16, np.asarray(samples.cpu()),
                                                                  np.linspace(-1, 1, m), np.linspace(-1, 1, m)), 20.0)
        cores = len(os.sched_getaffinity(0))
        cpu = {"value": m * m / t, "unit": "points/s", "cores": cores, "kind": "port",
               "sample": f"oracle interp_from_interior_2D to a {m}x{m} grid (median of {reps}); "
                         "the oracle's to-interior direction is not timed"}
    return _line("HPS<->grid interpolation points/s (both directions)", "points/s", npts / (ms / 1e3), ms,
                 args.steps, args.warmup, "f64",
                 {"workload": "interpolation 2D p=16 L=7: 64x64 Chebyshev samples -> HPS grid -> 1024x1024 grid",
                  "n_leaves": dom.n_leaves, "grid": ngrid, "samples": nsamp}, roof, kernels, cpu)


def adaptive_workload(p=16, q=14, tol=1e-12, nbumps=16):
    """Adaptive 2D Poisson problem: a tree refined around nbumps seeded sharp
    Gaussian bumps (1468 leaves, depth 7 at the defaults), source and
    boundary data of their sum."""
    import numpy as np

    import jaxhps_amd as jh

    rng = np.random.default_rng(0)
    cs = list(zip(rng.uniform(-0.8, 0.8, nbumps), rng.uniform(-0.8, 0.8, nbumps), rng.uniform(1000, 5000, nbumps)))

    def u(x):
        return sum(np.exp(-a * ((x[..., 0] - cx) ** 2 + (x[..., 1] - cy) ** 2)) for cx, cy, a in cs)

    def lap(x):
        out = 0.0
        for cx, cy, a in cs:
            r2 = (x[..., 0] - cx) ** 2 + (x[..., 1] - cy) ** 2
            out = out + (4 * a * a * r2 - 4 * a) * np.exp(-a * r2)
        return out

    dom = jh.Domain.from_adaptive_discretization(p, q, jh.DiscretizationNode2D(-1.0, 1.0, -1.0, 1.0), u, tol)
    return dom, u, lap


def bench_adaptive(args):
    """Adaptive 2D DtN build + solve (FP64).  Step = build_solver + solve."""
    import numpy as np
    import torch

    import jaxhps_amd as jh
    from jaxhps_amd import profiling

    dev = torch.device("cuda", 0)
    dom, u, lap = adaptive_workload()
    pts = dom.interior_points
    ones = torch.ones(pts.shape[:2], dtype=torch.float64, device=dev)
    f = torch.as_tensor(lap(pts)).to(dev)
    g = torch.as_tensor(np.concatenate(dom.get_adaptive_boundary_data_lst(u))).to(dev)
    pde = jh.PDEProblem(dom, source=f, D_xx_coefficients=ones, D_yy_coefficients=ones)

    def step():
        jh.build_solver(pde)
        return jh.solve(pde, g)

    ms = _timed_steps(step, args.steps, args.warmup)
    sol, kernels, roof = profiling.instrumented(step)
    err = float(np.abs(sol.cpu().numpy() - u(pts)).max())
    dof = dom.n_leaves * dom.p**2
    cpu = None
    if not args.no_cpu_baseline:
        from oracle import adaptive2d as ad

        one = np.ones(pts.shape[:2])
        fh, gh = lap(pts), g.cpu().numpy()
        t, reps = _time_bounded(lambda: ad.solve(dom.root, dom.p, dom.q, {"D_xx": one, "D_yy": one}, fh, gh), 20.0)
        cores = len(os.sched_getaffinity(0))
        cpu = {"value": dof / t, "unit": "DOF/s", "cores": cores, "kind": "port",
               "sample": f"adaptive2d oracle, the full build+solve of this tree (median of {reps})"}
    out = _line("adaptive 2D DtN build+solve DOF/s", "DOF/s", dof / (ms / 1e3), ms, args.steps, args.warmup, "f64",
                {"workload": "adaptive 2D Poisson, 16 Gaussian bumps, p=16 q=14 tol=1e-12",
                 "n_leaves": dom.n_leaves, "max_depth": int(max(lf.depth for lf in jh.get_all_leaves(dom.root)))},
                roof, kernels, cpu)
    out["max_abs_err_vs_exact"] = err
    return out


def bench_adaptive3d(args, p=8, q=6, tol=1e-6):
    """Adaptive 3D DtN build + solve (FP64) of a Poisson problem refined
    around four seeded Gaussian bumps (1331 leaves, depth 5 at the
    defaults).  Step = build_solver + solve."""
    import numpy as np
    import torch

    import jaxhps_amd as jh
    from jaxhps_amd import profiling

    dev = torch.device("cuda", 0)
    rng = np.random.default_rng(0)
    cs = list(zip(rng.uniform(0.2, 0.8, 4), rng.uniform(0.2, 0.8, 4), rng.uniform(0.2, 0.8, 4),
                  rng.uniform(100, 300, 4)))

    def u(x):
        return sum(np.exp(-a * ((x[..., 0] - cx) ** 2 + (x[..., 1] - cy) ** 2 + (x[..., 2] - cz) ** 2))
                   for cx, cy, cz, a in cs)

    def lap(x):
        out = 0.0
        for cx, cy, cz, a in cs:
            r2 = (x[..., 0] - cx) ** 2 + (x[..., 1] - cy) ** 2 + (x[..., 2] - cz) ** 2
            out = out + (4 * a * a * r2 - 6 * a) * np.exp(-a * r2)
        return out

    dom = jh.Domain.from_adaptive_discretization(p, q, jh.DiscretizationNode3D(0.0, 1.0, 0.0, 1.0, 0.0, 1.0), u, tol)
    pts = dom.interior_points
    ones = torch.ones(pts.shape[:2], dtype=torch.float64, device=dev)
    f = torch.as_tensor(lap(pts)).to(dev)
    g = torch.as_tensor(np.concatenate(dom.get_adaptive_boundary_data_lst(u))).to(dev)
    pde = jh.PDEProblem(dom, source=f, D_xx_coefficients=ones, D_yy_coefficients=ones, D_zz_coefficients=ones)

    def step():
        jh.build_solver(pde)
        return jh.solve(pde, g)

    ms = _timed_steps(step, args.steps, args.warmup)
    sol, kernels, roof = profiling.instrumented(step)
    err = float(np.abs(sol.cpu().numpy() - u(pts)).max())
    dof = dom.n_leaves * p**3
    cpu = None
    if not args.no_cpu_baseline:
        from oracle import adaptive3d as ad

        o = ad.Node((0.0, 1.0, 0.0, 1.0, 0.0, 1.0))

        def mirror(nd, on):
            if len(nd.children):
                ad.split(on)
                for c, oc in zip(nd.children, on.children):
                    mirror(c, oc)

        mirror(dom.root, o)
        one = np.ones(pts.shape[:2])
        fh, gh = lap(pts), g.cpu().numpy()
        t, reps = _time_bounded(lambda: ad.solve(o, p, q, {"D_xx": one, "D_yy": one, "D_zz": one}, fh, gh), 20.0)
        cores = len(os.sched_getaffinity(0))
        cpu = {"value": dof / t, "unit": "DOF/s", "cores": cores, "kind": "port",
               "sample": f"adaptive3d oracle, the full build+solve of this tree (median of {reps})"}
    out = _line("adaptive 3D DtN build+solve DOF/s", "DOF/s", dof / (ms / 1e3), ms, args.steps, args.warmup, "f64",
                {"workload": f"adaptive 3D Poisson, 4 Gaussian bumps, p={p} q={q} tol={tol:g}",
                 "n_leaves": dom.n_leaves, "max_depth": int(max(lf.depth for lf in jh.get_all_leaves(dom.root)))},
                roof, kernels, cpu)
    out["max_abs_err_vs_exact"] = err
    return out


def bench_subtree(args, L=8, height=6):
    """Subtree recomputation at a depth past the headline: the 2D Helmholtz
    workload (k = 100, p = 16) at L = 8 (65,536 leaves, 16.8M DOF), solved by
    solve_subtree (subtrees of height 6 rebuilt on the way down, nothing below
    their roots kept) and by build_solver + solve (every factor kept), FP64.
    Step = one solve_subtree; the line records both paths' times, their peak
    device memory and their relative difference."""
    import numpy as np
    import torch

    import jaxhps_amd as jh
    from jaxhps_amd import workloads as wl

    dev = torch.device("cuda", 0)
    dom, kw, g = wl.config_C(L=L)
    dkw = {k: (torch.as_tensor(v).to(dev) if isinstance(v, np.ndarray) else v) for k, v in kw.items()}
    g_dev = torch.as_tensor(g).to(dev)
    dof = dom.n_leaves * dom.p**2
    out = {}
    sols = {}
    for name in ("subtree", "stored"):
        pde = jh.PDEProblem(dom, **dkw)

        def step():
            if name == "subtree":
                return jh.solve_subtree(pde, g_dev, subtree_height=height)
            jh.build_solver(pde)
            return jh.solve(pde, g_dev)

        torch.cuda.synchronize()
        torch.cuda.empty_cache()
        torch.cuda.reset_peak_memory_stats(dev)
        base = torch.cuda.memory_allocated(dev)
        ms = _timed_steps(step, max(1, args.steps), args.warmup)
        sols[name] = step()
        torch.cuda.synchronize()
        out[name] = {"ms_per_step": ms, "value": dof / (ms / 1e3),
                     "peak_device_bytes": int(torch.cuda.max_memory_allocated(dev) - base)}
        del pde
        print(f"subtree bench: {name} {ms:.1f} ms", file=sys.stderr, flush=True)
    rel = float(torch.linalg.vector_norm(sols["subtree"] - sols["stored"]) / torch.linalg.vector_norm(sols["stored"]))
    line = _line(f"solve_subtree DOF/s at L={L} (subtree height {height})", "DOF/s", out["subtree"]["value"],
                 out["subtree"]["ms_per_step"], max(1, args.steps), args.warmup, "c128",
                 {"workload": "2D Helmholtz ItI k=100 p=16", "L": L, "n_leaves": dom.n_leaves,
                  "subtree_height": height}, None, None, None)
    line["paths"] = out
    line["rel_diff_subtree_vs_stored"] = rel
    line["accuracy_ok"] = rel < 1e-8  # the same discrete solution by two routes
    line["memory_ratio_stored_over_subtree"] = out["stored"]["peak_device_bytes"] / out["subtree"]["peak_device_bytes"]
    return line


def main():
    args = parse()
    if args.gpus != 1:
        print(
            f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported "
            "(the multi-GPU design is documented, not built)",
            file=sys.stderr,
        )
        sys.exit(2)
    import numpy as np
    import torch

    import jaxhps_amd as jh
    from jaxhps_amd import _lib, profiling

    if args.copy_calibration:
        x = torch.zeros(2**24, dtype=torch.complex128, device="cuda:0")  # 256 MiB
        lib = _lib.load_library()
        y = torch.empty_like(x)
        for _ in range(10):
            rc = lib.hps_convert(_lib.handle(), 2, 2, x.numel(), _lib.ptr(x), _lib.ptr(y))
            _lib.check(rc, _lib.handle(), "hps_convert")
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration": {"kernel": "convert_kernel<c128,c128>", "launches": 10,
                                               "read_bytes_per_launch": x.numel() * 16,
                                               "write_bytes_per_launch": x.numel() * 16}}), flush=True)
        return

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    if args.workload != "hps":
        fn = {"scattering": bench_scattering, "interp": bench_interp, "adaptive": bench_adaptive,
              "adaptive3d": bench_adaptive3d, "subtree": bench_subtree}[args.workload]
        print(json.dumps(fn(args)), flush=True)
        return
    dom, kw, g, desc = make_workload(args.config, args.L)
    # inputs resident in HBM before the timed region
    dkw = {}
    for k, v in kw.items():
        if isinstance(v, np.ndarray):
            dkw[k] = torch.as_tensor(v).to(dev)
        else:
            dkw[k] = v
    g_dev = torch.as_tensor(g).to(dev)
    dof = dom.n_leaves * dom.p ** (2 if dom.bool_2D else 3)
    paths = ["mixed", "fp64"] if args.precision == "both" else [args.precision]
    if args.config == "A" and args.precision == "both":
        paths = ["fp64"]  # the plain-method configuration
    low_leaves = {"auto": None, "low": True, "high": False}[args.leaf_precision]

    def build_kwargs(precision, top=None):
        if precision == "fp64":
            return {}
        return dict(precision="mixed", fp64_top_levels=args.fp64_top_levels if top is None else top,
                    low_precision_leaves=low_leaves)

    def timed(precision, top=None):
        pde = jh.PDEProblem(dom, **dkw)
        bkw = build_kwargs(precision, top)

        def step():
            jh.build_solver(pde, **bkw)
            if precision == "mixed":
                return jh.solve(pde, g_dev, refine_tol=args.refine_tol)
            return jh.solve(pde, g_dev)

        for _ in range(args.warmup):
            step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            u = step()
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) / args.steps * 1e3
        return pde, u, ms

    res = {}
    for prec in paths:
        pde, u, ms = timed(prec)
        res[prec] = dict(pde=pde, u=u, ms=ms, hist=list(getattr(pde, "refine_history", [])),
                         converged=getattr(pde, "refine_converged", None),
                         leaf_dtype=str(pde._factors["leaf"].lu.dtype).replace("torch.", ""),
                         level_dtypes=[str(S.dtype).replace("torch.", "") for S in pde.S_lst])
    head = res[paths[0]]
    ms = head["ms"]
    is_c = args.config == "C"
    dtypes = sorted({head["leaf_dtype"], *head["level_dtypes"]}, key=lambda d: ("64" in d, d))
    short = {"complex64": "c64", "complex128": "c128", "float32": "f32", "float64": "f64"}
    out = {
        "metric": "build+solve DOF/s, mixed vs FP64 GPU at <=1e-8 rel err; adjoint grads/s",
        "value": dof / (ms / 1e3),
        "unit": "DOF/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "+".join(short[d] for d in dtypes) + ("+c128 refine" if paths[0] == "mixed" and is_c else
                                                       "+f64 refine" if paths[0] == "mixed" else ""),
        "data": "synthetic (seeded coefficient fields and boundary data of the config)",
        "config": dict(desc, path=paths[0], fp64_top_levels=args.fp64_top_levels if paths[0] == "mixed" else None,
                       leaf_dtype=head["leaf_dtype"], merge_level_dtypes=head["level_dtypes"]),
    }
    if "mixed" in res:
        out["refine"] = {"tol": args.refine_tol, "iterations": len(res["mixed"]["hist"]),
                         "converged": res["mixed"]["converged"], "history": res["mixed"]["hist"]}
    if "fp64" in res and "mixed" in res:
        um, u64 = res["mixed"]["u"], res["fp64"]["u"]
        rel = float(torch.linalg.vector_norm(um - u64) / torch.linalg.vector_norm(u64))
        out["fp64_path"] = {"ms_per_step": res["fp64"]["ms"], "value": dof / (res["fp64"]["ms"] / 1e3)}
        out["speedup_vs_fp64"] = res["fp64"]["ms"] / res["mixed"]["ms"]
        out["rel_err_vs_fp64"] = rel
        if not args.no_variants and dom.L >= 2:
            pde2, u2, ms2 = timed("mixed", top=2)
            out["variant_fp64_top2"] = {
                "ms_per_step": ms2, "value": dof / (ms2 / 1e3), "speedup_vs_fp64": res["fp64"]["ms"] / ms2,
                "rel_err_vs_fp64": float(torch.linalg.vector_norm(u2 - u64) / torch.linalg.vector_norm(u64)),
                "refine_iterations": len(pde2.refine_history), "converged": pde2.refine_converged}
            del pde2, u2
    if not args.profile_only:
        # per-stage breakdown and dominant-kernel roofline from one instrumented step
        pde = head["pde"]
        if paths[0] == "mixed":
            prof = profiling.instrumented_step(
                pde, g_dev,
                build=lambda p: jh.build_solver(p, **build_kwargs("mixed")),
                solve=lambda p, gg: jh.solve(p, gg, refine_tol=args.refine_tol))
        else:
            prof = profiling.instrumented_step(pde, g_dev)
        out["stages_ms"] = prof["stages_ms"]
        out["roofline"] = prof["roofline"]
        if out["roofline"] is not None:
            traffic, src = measured_traffic(out["roofline"]["kernel"])
            out["roofline"]["traffic"] = traffic
            out["roofline"]["traffic_source"] = src
        out["kernels"] = prof["kernels"]
        if "fp64" in res and paths[0] != "fp64":
            p64 = profiling.instrumented_step(res["fp64"]["pde"], g_dev)
            out["fp64_path"]["stages_ms"] = p64["stages_ms"]
            out["fp64_path"]["roofline"] = p64["roofline"]
    if not args.no_cpu_baseline and not args.profile_only:
        # the CPU leg: the oracle timed on the host cores, and (2D) both GPU
        # paths checked against it at L = 4 of the same workload
        out["cpu_baseline"] = cpu_baseline(args.config)
        if dom.bool_2D and not args.no_oracle_check:
            out["oracle_check"] = oracle_check(args.config, build_kwargs("mixed"), args.refine_tol)
    converged = bool(res["mixed"]["converged"]) if "mixed" in res else True
    if args.config == "C" and not args.profile_only and not (args.no_adjoint and args.no_other_configs):
        for prec in res.values():
            prec.pop("pde", None)
        res.clear()
        torch.cuda.empty_cache()
    if args.config == "C" and not args.no_other_configs and not args.profile_only:
        out["other_configs"] = other_configs_bench(args.refine_tol)
        torch.cuda.empty_cache()
    if args.config == "C" and not args.no_adjoint and not args.profile_only:
        out["adjoint"] = adjoint_bench(args.adjoint_L)
        if not args.no_surrogate:
            torch.cuda.empty_cache()
            out["adjoint"]["surrogate_loop"] = surrogate_loop_bench(args.adjoint_L)
    ok = converged and out.get("rel_err_vs_fp64", 0.0) <= 1e-8
    oc = out.get("oracle_check")
    if oc:
        ok = ok and oc["fp64_vs_oracle"] <= 1e-10 and oc["fp64_vs_exact"] <= 1e-10
        ok = ok and oc["mixed_vs_oracle"] <= 1e-8 and oc["mixed_vs_exact"] <= 1e-8
    out["accuracy_ok"] = ok
    print(json.dumps(out), flush=True)
    if not ok:
        print("bench.py: the accuracy bar (mixed <= 1e-8 vs FP64 and the oracle, FP64 <= 1e-10 vs the oracle, "
              "refinement converged) is not met; the value above is not valid", file=sys.stderr)
        sys.exit(3)


if __name__ == "__main__":
    main()
