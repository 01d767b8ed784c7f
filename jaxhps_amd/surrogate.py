"""Learned correction to the leaf ItI map (no counterpart in the reference).

A 3-layer MLP (tanh, width 512) maps a leaf's 256 coefficient samples q to a
rank-r complex correction of the q = 0 leaf ItI map T0:

    T(q) ~= T0 + s U V^H,      U, V in C^{nq x r}   (nq = 56, r = 8 for p = 16)

It is trained on leaf solves the method itself generates (leaf_training_data
runs the ordinary FP64 leaf stage on random coefficient fields of the
inverse-scattering family, q = sum_j theta_j sine_j), with hand-written
kernels (libhps_mi355x ``hps_mlp_*``): BF16 MFMA products with FP32
accumulation for the forward and backward passes, tanh/bias/Adam kernels,
and the low-rank loss |U V^H - (T - T0)/s|_F^2 with its gradient.

surrogate_leaf_maps checks every surrogate leaf against the plain FP64 leaf
solve on a random boundary probe and falls back to the plain map where the
probe's relative error exceeds the tolerance; the plain method judges the
end-to-end error.
"""

from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from ._lib import check, load_library


def _f(t):
    return _lib.ptr(t)


class LeafItISurrogate:
    """MLP n_in -> width -> width -> 4 nq r with FP32 master weights on the GPU."""

    def __init__(self, n_in=256, width=512, nq=56, rank=8, seed=0, device=None):
        self.device = torch.device(device) if device is not None else torch.device("cuda", 0)
        self.n_in, self.width, self.nq, self.rank = n_in, width, nq, rank
        self.n_out = 4 * nq * rank
        g = torch.Generator().manual_seed(seed)

        def init(o, i, gain=1.0):
            return (torch.randn((o, i), generator=g, dtype=torch.float32) * gain / math.sqrt(i)).to(self.device)

        self.params = {
            "W1": init(width, n_in), "b1": torch.zeros(width, device=self.device),
            "W2": init(width, width), "b2": torch.zeros(width, device=self.device),
            "W3": init(self.n_out, width, 0.1), "b3": torch.zeros(self.n_out, device=self.device),
        }
        self.grads = {k: torch.zeros_like(v) for k, v in self.params.items()}
        self.m1 = {k: torch.zeros_like(v) for k, v in self.params.items()}
        self.m2 = {k: torch.zeros_like(v) for k, v in self.params.items()}
        self.step = 0
        self.in_shift, self.in_scale, self.out_scale = 0.0, 1.0, 1.0
        self.T0 = None  # (nq, nq) complex128 on the device

    # -- kernels -------------------------------------------------------------
    def _gemm(self, ta, tb, M, N, K, A, lda, B, ldb, C, ldc, bias=None, act=0, beta=0.0):
        lib = load_library()
        h = _lib.handle()
        rc = lib.hps_mlp_gemm_bf16(h, ta, tb, M, N, K, _f(A), lda, _f(B), ldb, beta, _f(C), ldc,
                                   _f(bias) if bias is not None else None, act)
        check(rc, h, "hps_mlp_gemm_bf16")

    def forward(self, X):
        """X (batch, n_in) float32 (normalised) -> (H1, H2, Y)."""
        p = self.params
        B = X.shape[0]
        H1 = torch.empty((B, self.width), dtype=torch.float32, device=self.device)
        H2 = torch.empty_like(H1)
        Y = torch.empty((B, self.n_out), dtype=torch.float32, device=self.device)
        self._gemm(0, 1, B, self.width, self.n_in, X, self.n_in, p["W1"], self.n_in, H1, self.width, p["b1"], 1)
        self._gemm(0, 1, B, self.width, self.width, H1, self.width, p["W2"], self.width, H2, self.width, p["b2"], 1)
        self._gemm(0, 1, B, self.n_out, self.width, H2, self.width, p["W3"], self.width, Y, self.n_out, p["b3"], 0)
        return H1, H2, Y

    def backward(self, X, H1, H2, dY):
        """Parameter gradients into self.grads from dY (batch, n_out)."""
        lib = load_library()
        h = _lib.handle()
        p, gr = self.params, self.grads
        B = X.shape[0]
        w, o = self.width, self.n_out
        self._gemm(1, 0, o, w, B, dY, o, H2, w, gr["W3"], w)
        check(lib.hps_mlp_bias_grad(h, B, o, _f(dY), _f(gr["b3"])), h, "bias_grad")
        dH2 = torch.empty((B, w), dtype=torch.float32, device=self.device)
        self._gemm(0, 0, B, w, o, dY, o, p["W3"], w, dH2, w)
        check(lib.hps_mlp_tanh_backward(h, dH2.numel(), _f(H2), _f(dH2)), h, "tanh_backward")
        self._gemm(1, 0, w, w, B, dH2, w, H1, w, gr["W2"], w)
        check(lib.hps_mlp_bias_grad(h, B, w, _f(dH2), _f(gr["b2"])), h, "bias_grad")
        dH1 = torch.empty((B, w), dtype=torch.float32, device=self.device)
        self._gemm(0, 0, B, w, w, dH2, w, p["W2"], w, dH1, w)
        check(lib.hps_mlp_tanh_backward(h, dH1.numel(), _f(H1), _f(dH1)), h, "tanh_backward")
        self._gemm(1, 0, w, self.n_in, B, dH1, w, X, self.n_in, gr["W1"], self.n_in)
        check(lib.hps_mlp_bias_grad(h, B, w, _f(dH1), _f(gr["b1"])), h, "bias_grad")

    def loss_and_grad(self, Y, dT_scaled):
        """Mean over samples of |U V^H - dT_scaled|_F^2 and dL/dY."""
        lib = load_library()
        h = _lib.handle()
        B = Y.shape[0]
        dY = torch.empty_like(Y)
        loss = torch.empty(B, dtype=torch.float32, device=self.device)
        dTf = torch.view_as_real(dT_scaled.contiguous()).contiguous()
        rc = lib.hps_mlp_lowrank_loss(h, B, self.nq, self.rank, _f(Y), _f(dTf), 1.0 / B, _f(dY), _f(loss))
        check(rc, h, "hps_mlp_lowrank_loss")
        return loss.sum(), dY

    def adam_step(self, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
        lib = load_library()
        h = _lib.handle()
        self.step += 1
        for k, w in self.params.items():
            rc = lib.hps_mlp_adam(h, w.numel(), _f(w), _f(self.grads[k]), _f(self.m1[k]), _f(self.m2[k]), lr, b1,
                                  b2, eps, self.step)
            check(rc, h, "hps_mlp_adam")

    # -- training / inference ----------------------------------------------
    def normalise(self, q):
        return ((q - self.in_shift) / self.in_scale).to(torch.float32).contiguous()

    def fit(self, q, T, T0, steps=1500, batch=256, lr=1e-3, seed=0, log_every=0):
        """Train on leaf samples q (n, n_in) with exact maps T (n, nq, nq) complex."""
        q = q.to(self.device, torch.float64)
        T = T.to(self.device, torch.complex128)
        self.T0 = T0.to(self.device, torch.complex128).contiguous()
        dT = T - self.T0
        self.in_shift = float(q.mean())
        self.in_scale = float(q.std()) or 1.0
        self.out_scale = float(dT.abs().pow(2).mean().sqrt()) or 1.0
        X = self.normalise(q)
        target = (dT / self.out_scale).to(torch.complex64)
        g = torch.Generator(device="cpu").manual_seed(seed)
        hist = []
        for it in range(steps):
            idx = torch.randint(0, X.shape[0], (batch,), generator=g).to(self.device)
            xb = X[idx].contiguous()
            H1, H2, Y = self.forward(xb)
            loss, dY = self.loss_and_grad(Y, target[idx])
            self.backward(xb, H1, H2, dY)
            cur_lr = lr * 0.5 * (1 + math.cos(math.pi * it / steps))
            self.adam_step(cur_lr)
            if log_every and (it % log_every == 0 or it == steps - 1):
                hist.append((it, float(loss)))
        return hist

    def predict_T(self, q):
        """Surrogate maps (n, nq, nq) complex128 for leaf samples q (n, n_in)."""
        lib = load_library()
        h = _lib.handle()
        X = self.normalise(q.to(self.device, torch.float64))
        _, _, Y = self.forward(X)
        out = torch.empty((X.shape[0], self.nq, self.nq), dtype=torch.complex128, device=self.device)
        rc = lib.hps_mlp_lowrank_apply(h, X.shape[0], self.nq, self.rank, _f(Y),
                                       _f(torch.view_as_real(self.T0)), self.out_scale, _f(torch.view_as_real(out)))
        check(rc, h, "hps_mlp_lowrank_apply")
        return out


def leaf_training_data(n_fields=8, L=6, p=16, k=40.0, sigma=0.1, n_terms=25, seed=0, device=None):
    """(q (n, p^2), T (n, nq, nq), T0 (nq, nq)) from the FP64 leaf stage on
    random fields q = sum_j theta_j sine_j, theta ~ N(0, sigma^2)."""
    from . import workloads as wl
    from ._pdeproblem import PDEProblem
    from .local_solve import local_solve_stage_uniform_2D_ItI

    dev = torch.device(device) if device is not None else torch.device("cuda", 0)
    dom, kw, _, basis, _, _ = wl.config_E(L=L, p=p, k=k, n_theta=1, n_terms=n_terms, seed=seed)
    rng = np.random.default_rng(seed + 1)
    qs, Ts = [], []
    for _ in range(n_fields):
        theta = sigma * rng.standard_normal(n_terms)
        q = np.tensordot(theta, basis, axes=1)
        pde = PDEProblem(dom, I_coefficients=k**2 * (1.0 - q), **kw)
        _, T, _, _ = local_solve_stage_uniform_2D_ItI(pde, device=dev)
        qs.append(torch.as_tensor(q, device=dev))
        Ts.append(T)
    pde0 = PDEProblem(dom, I_coefficients=k**2 * np.ones(dom.interior_points.shape[:2]), **kw)
    _, T0, _, _ = local_solve_stage_uniform_2D_ItI(pde0, device=dev)
    return torch.cat(qs), torch.cat(Ts), T0[0]


@dataclass
class SurrogateCheck:
    T: torch.Tensor           # (n_leaves, nq, nq) maps used: surrogate where it passed, plain elsewhere
    passed: torch.Tensor      # (n_leaves,) bool
    probe_error: torch.Tensor  # (n_leaves,) relative error of the surrogate on the probe


def surrogate_leaf_maps(model: LeafItISurrogate, q, T_plain, tol=1e-2, seed=0):
    """Check every surrogate leaf against the plain leaf map on one random
    boundary probe; leaves whose relative probe error exceeds tol fall back to
    the plain map."""
    T_sur = model.predict_T(q)
    g = torch.Generator(device="cpu").manual_seed(seed)
    probe = torch.complex(torch.randn((q.shape[0], model.nq, 1), generator=g, dtype=torch.float64),
                          torch.randn((q.shape[0], model.nq, 1), generator=g, dtype=torch.float64)).to(T_sur.device)
    t_s = _lib.gemm(T_sur, probe)
    t_p = _lib.gemm(T_plain.contiguous(), probe)
    err = (t_s - t_p).abs().pow(2).sum((1, 2)).sqrt() / t_p.abs().pow(2).sum((1, 2)).sqrt()
    passed = err <= tol
    T = torch.where(passed[:, None, None], T_sur, T_plain)
    return SurrogateCheck(T, passed, err)


# ---------------------------------------------------------------------------
# The surrogate inside an optimisation loop
# ---------------------------------------------------------------------------
class SharedLeafSolver:
    """Solves with every ItI leaf matrix of a uniform 2D tree whose leaves
    differ only in the identity coefficient (the inverse-scattering family
    I = k^2 (1 - q)): B(q) = B_ref + diag(0, d) with d = I - I_ref on the
    collocation rows.  B_ref (the same matrix for every leaf: the tree is
    uniform) is factored once; B(q) x = r is solved by the fixed point
    x = B_ref^{-1} (r - d x), all leaves and right-hand sides as the columns
    of one product with the explicit inverse of B_ref per step (formed once
    from its LU; B_ref is row-scaled, condition ~1e4, so the fixed point
    differs from B(q)^{-1} r by ~1e-12 relative).  It contracts by about
    |B_ref^{-1} diag(d)| (about 0.1 for k = 40, p = 16, L = 6, |q| <= 0.5);
    leaves that do not converge are reported."""

    def __init__(self, pde, I_ref, device=None):
        from ._stages import default_device

        dom = pde.domain
        if not (pde.use_ItI and dom.bool_2D and dom.bool_uniform):
            raise ValueError("SharedLeafSolver needs a uniform 2D ItI problem")
        self.device = torch.device(device) if device is not None else default_device()
        n = dom.p**2
        A = np.zeros((n, n))
        for nm, arr in pde.coefficient_list():
            if nm == "I":
                continue
            a = np.asarray(arr.detach().cpu() if isinstance(arr, torch.Tensor) else arr)
            if not np.all(a == a.flat[0]):
                raise ValueError(f"{nm} coefficients vary: the leaf matrices differ beyond the identity term")
            A += float(np.real(a.flat[0])) * np.asarray(pde_op(pde, nm))
        A += float(I_ref) * np.eye(n)
        G = np.asarray(pde.G)
        self.nb = G.shape[0]
        self.s = float(2.0 ** np.round(np.log2(np.abs(A[self.nb:]).max() / np.abs(G).max())))
        B = np.concatenate([self.s * G, A[self.nb:].astype(complex)])
        LU = torch.as_tensor(B[None], dtype=torch.complex128).to(self.device).contiguous()
        perm, dinv, info = _lib.getrf(LU)
        _lib.raise_on_info(info, "shared leaf matrix")
        eye = torch.eye(n, dtype=torch.complex128, device=self.device)[None].contiguous()
        self.Binv = _lib.getrs(LU, perm, dinv, eye)[0].contiguous()
        self.I_ref = float(I_ref)
        self.n = n

    def _apply(self, R):
        """B_ref^{-1} applied to every leaf's columns of R (nl, n, w)."""
        nl, n, w = R.shape
        Rt = R.permute(1, 0, 2).reshape(n, nl * w).contiguous()
        X = _lib.gemm(self.Binv, Rt)[0]
        return X.reshape(n, nl, w).permute(1, 0, 2).contiguous()

    def solve(self, rhs, I_coef, tol=1e-12, max_iter=60, chunk=4):
        """x (nl, n, w) with B(q) x = rhs (impedance rows first, unscaled).
        Returns (x, converged (nl,) bool, iterations)."""
        nl, n, w = rhs.shape
        d = (I_coef.to(self.device, torch.float64) - self.I_ref).to(torch.complex128)
        d = d.clone()
        d[:, : self.nb] = 0
        d = d.unsqueeze(-1)
        r = rhs.to(torch.complex128).clone()
        r[:, : self.nb] *= self.s
        x = self._apply(r)
        it = 0
        rel = None
        while it < max_iter:
            for _ in range(chunk):
                x_new = self._apply(r - d * x)
                it += 1
                diff = (x_new - x).abs().pow(2).sum((1, 2))
                x = x_new
            rel = (diff / x.abs().pow(2).sum((1, 2)).clamp_min(1e-300)).sqrt()
            if float(rel.max()) <= tol:
                break
        return x, torch.isfinite(rel) & (rel <= max(tol, 1e-10)), it


def pde_op(pde, name):
    return getattr(pde, name)


@dataclass
class SurrogateSolve:
    u: torch.Tensor            # (n_leaves, n) complex128
    passed: torch.Tensor       # (n_leaves,) bool: the surrogate map was used
    probe_error: torch.Tensor  # (n_leaves,) relative error of the surrogate map on the probe
    iterations: int            # fixed-point steps of the leaf solves (probe + interior)


def surrogate_solve(model: LeafItISurrogate, solver: SharedLeafSolver, pde, boundary_data, I_coef, k,
                    tol=1e-2, merge_dtype=torch.complex64, seed=0, leaf_tol=1e-9) -> SurrogateSolve:
    """u for the Helmholtz problem with I = I_coef (n_leaves, n) = k^2 (1 - q)
    and zero source, with the leaf ItI maps from the surrogate.

    Every surrogate leaf map is checked on a random boundary probe t against
    the plain leaf solve of the same probe (B(q) x = [P t; 0], T t = QH x,
    solved with the shared factors); leaves whose relative probe error exceeds
    tol (or whose solve does not converge) fall back to the plain leaf stage.
    The merges run in merge_dtype; each leaf's interior comes from its plain
    solve with the incoming impedance of the down pass.  The plain method
    judges the end-to-end error (compare with build_solver + solve)."""
    from . import _tables
    from ._stages import leaf_solve, to_device
    from .down_pass import down_levels
    from .merge import _levels

    dom = pde.domain
    dev = solver.device
    I_t = to_device(I_coef, dev, torch.float64)
    q = 1.0 - I_t / k**2
    nl, n = I_t.shape
    ops = pde.device_ops(dev, torch.complex128)
    P, QH = ops["P"], ops["Q"]
    nb, nq = P.shape
    T = model.predict_T(q)
    g = torch.Generator(device="cpu").manual_seed(seed)
    t = torch.complex(torch.randn((nl, nq, 1), generator=g, dtype=torch.float64),
                      torch.randn((nl, nq, 1), generator=g, dtype=torch.float64)).to(dev)
    rhs = torch.zeros((nl, n, 1), dtype=torch.complex128, device=dev)
    rhs[:, :nb] = _lib.gemm(P, t)
    x, conv, it1 = solver.solve(rhs, I_t, tol=leaf_tol)
    Tt = _lib.gemm(QH, x)
    err = ((_lib.gemm(T, t) - Tt).abs().pow(2).sum((1, 2)) / Tt.abs().pow(2).sum((1, 2))).sqrt()
    passed = conv & (err <= tol)
    fail = (~passed).nonzero().flatten()
    Y_fail = None
    if fail.numel():
        sub = _leaf_subproblem(pde, I_t, fail)
        res = leaf_solve(sub, 1, torch.complex128, dev, nsrc_zero=True)
        _lib.raise_on_info(res.info, "surrogate fallback leaves")
        T = T.clone()
        T[fail] = res.T
        Y_fail = res.Y
    L = dom.L
    levels, _ = _levels(T.to(merge_dtype).contiguous(), None, L, "iti", 4,
                        lambda nec: _tables.quad_merge_tables(nec // 4, "iti"), root_T=False)
    g_root = to_device(boundary_data, dev, merge_dtype).reshape(-1, 1)
    _, g_leaf = down_levels(g_root, [lv.S for lv in levels], [], None, None)
    g_leaf = g_leaf.to(torch.complex128)
    rhs = torch.zeros((nl, n, 1), dtype=torch.complex128, device=dev)
    rhs[:, :nb] = _lib.gemm(P, g_leaf)
    u, conv2, it2 = solver.solve(rhs, I_t, tol=leaf_tol)
    if fail.numel():
        u[fail] = _lib.gemm(Y_fail, g_leaf[fail].contiguous())
    return SurrogateSolve(u[..., 0], passed, err, it1 + it2)


def _leaf_subproblem(pde, I_t, idx):
    """The PDEProblem restricted to the leaves idx (same operators), with I
    coefficients I_t[idx]."""
    import copy

    from ._pdeproblem import PDEProblem

    sub = PDEProblem.__new__(PDEProblem)
    sub.__dict__.update({kk: v for kk, v in pde.__dict__.items() if not kk.endswith("_coefficients")})
    for nm in ("D_xx", "D_xy", "D_yy", "D_x", "D_y", "I"):
        a = getattr(pde, f"{nm}_coefficients", None)
        if nm == "I":
            a = I_t
        if a is not None:
            a = a[idx.cpu().numpy()] if not isinstance(a, torch.Tensor) else a[idx.to(a.device)]
        setattr(sub, f"{nm}_coefficients", a)
    sub.source = None
    dom = copy.copy(pde.domain)
    dom.n_leaves = int(idx.numel())
    sub.domain = dom
    sub._dev_cache = {kk: v for kk, v in pde._dev_cache.items() if not str(kk[0]).startswith(("refine", "iti_row"))}
    sub.reset()
    return sub
