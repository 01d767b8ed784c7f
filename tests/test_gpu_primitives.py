"""GPU parity of the batched dense primitives (GEMM, LU, LU solve) against
float64/complex128 references computed on the CPU."""

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

DTYPES = [torch.float64, torch.float32, torch.complex128, torch.complex64]
TOL = {torch.float64: 1e-12, torch.float32: 2e-5, torch.complex128: 1e-12, torch.complex64: 2e-5}


def _rand(shape, dtype, gen):
    x = torch.randn(shape, generator=gen, dtype=torch.float64)
    if dtype.is_complex:
        x = torch.complex(x, torch.randn(shape, generator=gen, dtype=torch.float64))
    return x.to(dtype)


def _ref64(t):
    return t.detach().cpu().to(torch.complex128 if t.is_complex() else torch.float64)


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize(
    "batch,M,N,K",
    [(1, 16, 16, 4), (3, 64, 64, 16), (5, 56, 56, 256), (2, 130, 70, 33), (7, 196, 72, 60), (2, 100, 1, 300),
     (3, 77, 16, 45), (4, 33, 3, 7), (1, 300, 257, 129), (6, 200, 2, 70), (3, 150, 9, 256),
     (3, 90, 1, 56), (2, 131, 2, 40), (2, 37, 2, 700), (1, 5, 1, 3), (1, 300, 4, 100), (2, 700, 3, 600),
     (1, 129, 4, 50)],
)
def test_gemm_matches_reference(gpu, dtype, batch, M, N, K):
    from jaxhps_amd import _lib

    print(f"gemm dtype={dtype} shape={(batch, M, N, K)}", flush=True)
    g = torch.Generator().manual_seed(batch * 1000 + M + N + K)
    A = _rand((batch, M, K), dtype, g)
    B = _rand((batch, K, N), dtype, g)
    C = _rand((batch, M, N), dtype, g)
    out = _lib.gemm(A.to(gpu), B.to(gpu), C.to(gpu), alpha=-0.5, beta=2.0)
    ref = -0.5 * (_ref64(A) @ _ref64(B)) + 2.0 * _ref64(C)
    err = (_ref64(out) - ref).abs().max().item() / max(1.0, ref.abs().max().item())
    assert err < TOL[dtype] * max(1, K) ** 0.5, err


@pytest.mark.parametrize("dtype", [torch.float64, torch.complex128])
def test_gemm_broadcast_operand(gpu, dtype):
    from jaxhps_amd import _lib

    g = torch.Generator().manual_seed(7)
    Q = _rand((56, 256), dtype, g)
    Y = _rand((9, 256, 56), dtype, g)
    out = _lib.gemm(Q.to(gpu), Y.to(gpu))
    ref = _ref64(Q) @ _ref64(Y)
    assert (_ref64(out) - ref).abs().max().item() < 1e-11 * ref.abs().max().item()


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize(
    "n,batch",
    [(5, 3), (36, 4), (196, 8), (256, 2), (33, 1), (700, 1), (300, 2), (1100, 1), (2500, 2), (4200, 1), (600, 110)],
)
def test_getrf_getrs_solve(gpu, dtype, n, batch):
    from jaxhps_amd import _lib

    print(f"lu dtype={dtype} n={n} batch={batch}", flush=True)
    g = torch.Generator().manual_seed(n + batch)
    A = _rand((batch, n, n), dtype, g)
    R = _rand((batch, n, 7), dtype, g)
    LU = A.to(gpu).contiguous()
    perm, dinv, info = _lib.getrf(LU)
    assert int(info.abs().sum()) == 0
    X = _lib.getrs(LU, perm, dinv, R.to(gpu).contiguous())
    ref = torch.linalg.solve(_ref64(A), _ref64(R))
    rel = ((_ref64(X) - ref).abs().max() / ref.abs().max()).item()
    cond_scale = n
    assert rel < TOL[dtype] * cond_scale * 50, rel
    # P A = L U reconstruction
    lu = _ref64(LU)
    L = torch.tril(lu, -1) + torch.eye(n, dtype=lu.dtype)
    U = torch.triu(lu)
    pa = torch.stack([_ref64(A)[b][perm[b].cpu().long()] for b in range(batch)])
    rec = ((L @ U - pa).abs().max() / _ref64(A).abs().max()).item()
    assert rec < TOL[dtype] * n * 10, rec


@pytest.mark.parametrize("dtype", [torch.complex128, torch.float32])
@pytest.mark.parametrize("n,batch", [(100, 3), (256, 2), (700, 1)])
def test_getrs_many_rhs(gpu, dtype, n, batch):
    from jaxhps_amd import _lib

    print(f"getrs many rhs dtype={dtype} n={n} batch={batch}", flush=True)
    g = torch.Generator().manual_seed(7 * n + batch)
    A = _rand((batch, n, n), dtype, g)
    R = _rand((batch, n, 45), dtype, g)
    LU = A.to(gpu).contiguous()
    perm, dinv, info = _lib.getrf(LU)
    X = _lib.getrs(LU, perm, dinv, R.to(gpu).contiguous())
    ref = torch.linalg.solve(_ref64(A), _ref64(R))
    rel = ((_ref64(X) - ref).abs().max() / ref.abs().max()).item()
    assert rel < TOL[dtype] * n * 50, rel


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("n,batch,nrhs", [(896, 4, 1), (520, 3, 16), (1100, 1, 2), (300, 127, 1), (300, 128, 3),
                                          (256, 5, 1), (300, 130, 1)])
def test_getrs_few_rhs(gpu, dtype, n, batch, nrhs):
    """Few right-hand sides: the blocked solve with the per-block kernel for
    small batches, the one-workgroup-per-matrix solve from 128 matrices (or
    n <= 256), with its single-column form."""
    from jaxhps_amd import _lib

    print(f"getrs few rhs dtype={dtype} n={n} batch={batch} nrhs={nrhs}", flush=True)
    g = torch.Generator().manual_seed(3 * n + batch + nrhs)
    A = _rand((batch, n, n), dtype, g)
    R = _rand((batch, n, nrhs), dtype, g)
    LU = A.to(gpu).contiguous()
    perm, dinv, info = _lib.getrf(LU)
    X = _lib.getrs(LU, perm, dinv, R.to(gpu).contiguous())
    ref = torch.linalg.solve(_ref64(A), _ref64(R))
    rel = ((_ref64(X) - ref).abs().max() / ref.abs().max()).item()
    assert rel < TOL[dtype] * n * 50, rel


def test_getrf_reports_singular(gpu):
    from jaxhps_amd import _lib

    A = torch.randn(3, 20, 20, dtype=torch.float64)
    A[1, :, 4] = 0.0  # column 4 of element 1 all zero
    LU = A.to(gpu).contiguous()
    perm, dinv, info = _lib.getrf(LU)
    info = info.cpu().tolist()
    assert info[0] == 0 and info[2] == 0
    assert info[1] == 5
    with pytest.raises(_lib.HPSSingularError):
        _lib.raise_on_info(torch.tensor(info), "test")


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize(
    "batch,M,N,K",
    [(3, 70, 1, 45), (2, 33, 2, 300), (4, 100, 9, 64), (2, 130, 70, 33), (2, 130, 70, 32), (3, 65, 97, 7),
     (2, 30, 70, 20)],
)
def test_gemm_conjt_matches_reference(gpu, dtype, batch, M, N, K):
    """D = alpha A^H B + beta C with A stored K x M."""
    from jaxhps_amd import _lib

    print(f"gemm_conjt dtype={dtype} shape={(batch, M, N, K)}", flush=True)
    g = torch.Generator().manual_seed(batch + M + N + K)
    A = _rand((batch, K, M), dtype, g)
    B = _rand((batch, K, N), dtype, g)
    C = _rand((batch, M, N), dtype, g)
    lib = _lib.load_library()
    h = _lib.handle()
    Ad, Bd, Cd = A.to(gpu), B.to(gpu), C.to(gpu)
    out = torch.empty((batch, M, N), dtype=dtype, device=gpu)
    rc = lib.hps_gemm_batched_conjt(h, _lib.dtype_code(dtype), M, N, K, 0.5, _lib.ptr(Ad), M, K * M, _lib.ptr(Bd), N,
                                    K * N, -1.0, _lib.ptr(Cd), N, M * N, _lib.ptr(out), N, M * N, batch)
    _lib.check(rc, h, "conjt")
    ref = 0.5 * (_ref64(A).conj().transpose(1, 2) @ _ref64(B)) - _ref64(C)
    err = (_ref64(out) - ref).abs().max().item() / max(1.0, ref.abs().max().item())
    assert err < TOL[dtype] * max(1, K) ** 0.5, err


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("n,nrhs,batch", [(37, 3, 4), (256, 1, 3), (1100, 5, 1)])
def test_getrs_conjt_solves_adjoint_system(gpu, dtype, n, nrhs, batch):
    from jaxhps_amd import _lib
    from jaxhps_amd.adjoint import getrs_conjt

    print(f"getrs_conjt dtype={dtype} n={n}", flush=True)
    g = torch.Generator().manual_seed(n + nrhs)
    A = _rand((batch, n, n), dtype, g) + n**0.5 * torch.eye(n, dtype=dtype)
    R = _rand((batch, n, nrhs), dtype, g)
    LU = A.to(gpu).contiguous()
    perm, dinv, info = _lib.getrf(LU)
    X = getrs_conjt(n, LU, 0, n, n * n, perm, dinv, R.to(gpu).contiguous())
    res = _ref64(A).conj().transpose(1, 2) @ _ref64(X) - _ref64(R)
    assert res.abs().max().item() < TOL[dtype] * 50 * n**0.5


@pytest.mark.parametrize("kind,dtype,m,nm", [("iti", torch.complex64, 1536, 1), ("iti", torch.complex64, 1536, 2),
                                             ("dtn", torch.float32, 768, 1)])
def test_block_inverse_solves_match_fp64_merge(gpu, kind, dtype, m, nm):
    """A merge level whose factored matrix has 3072 unknowns in low precision
    takes its many-column solve through the inverses of the 512 x 512
    diagonal blocks; S, T, g_tilde and h agree with the same merge in double
    precision (the standard blocked solve) to the low precision's accuracy."""
    from jaxhps_amd._stages import merge_level
    from jaxhps_amd._tables import quad_merge_tables

    ne = 4 * m
    print(f"block-inverse merge {kind} {dtype} m={m} matrices={nm}", flush=True)
    g = torch.Generator(device=gpu).manual_seed(5)
    X = torch.randn(4 * nm, ne, ne, generator=g, device=gpu, dtype=torch.float64) / ne**0.5
    if kind == "iti":
        Y = torch.randn(4 * nm, ne, ne, generator=g, device=gpu, dtype=torch.float64) / ne**0.5
        T64 = (0.5 * torch.complex(X, Y)).contiguous()
        high = torch.complex128
    else:
        T64 = (2.0 * torch.eye(ne, device=gpu, dtype=torch.float64) + 0.1 * X).contiguous()
        high = torch.float64
    del X
    h64 = torch.randn(4 * nm, ne, 2, generator=g, device=gpu, dtype=torch.float64).to(T64.dtype).contiguous()
    tb = quad_merge_tables(m, kind)
    lo = merge_level(T64.to(dtype).contiguous(), h64.to(dtype).contiguous(), tb)
    hi = merge_level(T64.to(high), h64.to(high), tb)
    nfac = lo.F_lu.shape[-1]
    assert nfac >= 3072 and nfac % 512 == 0
    for a, b in ((lo.S, hi.S), (lo.T, hi.T), (lo.gt, hi.gt), (lo.h, hi.h)):
        err = float((a.to(high) - b).abs().max() / b.abs().max())
        assert err < 2e-4, err


def test_unpivoted_panels_of_many_matrices(gpu):
    """96 complex64 interface matrices of 1200 unknowns: the leading panels
    (taller than 1024 rows) unpivoted, the rest partially pivoted, with more
    unpivoted-panel workgroups than the GPU holds at once, so the workgroups of
    one matrix do not all run together.  S, T, g_tilde and h equal the
    pivoted merge's (the factored diagonal block is written only after every
    workgroup of its matrix has read the original)."""
    from jaxhps_amd._stages import merge_factors_stable, merge_level
    from jaxhps_amd._tables import quad_merge_tables

    m, nm = 300, 96
    ne = 4 * m
    print(f"unpivoted panels of {nm} matrices, n = {ne}", flush=True)
    g = torch.Generator(device=gpu).manual_seed(11)
    X = torch.randn(4 * nm, ne, ne, generator=g, device=gpu, dtype=torch.float32) / ne**0.5
    Y = torch.randn(4 * nm, ne, ne, generator=g, device=gpu, dtype=torch.float32) / ne**0.5
    T = (0.5 * torch.complex(X, Y)).contiguous()
    del X, Y
    h = torch.randn(4 * nm, ne, 1, generator=g, device=gpu, dtype=torch.float32).to(torch.complex64).contiguous()
    tb = quad_merge_tables(m, "iti")
    ref = merge_level(T, h, tb, pivot=True)
    out = merge_level(T, h, tb, pivot=False)
    assert bool(merge_factors_stable(out.F_lu))
    for a, b in ((out.S, ref.S), (out.T, ref.T), (out.gt, ref.gt), (out.h, ref.h)):
        err = float((a - b).abs().amax(dim=(1, 2)).max() / b.abs().max())
        assert err < 2e-4, err


@pytest.mark.parametrize("kind,dtype", [("iti", torch.complex128), ("iti", torch.complex64),
                                        ("dtn", torch.float64), ("dtn", torch.float32)])
def test_unpivoted_tall_merge_matches_pivoted(gpu, kind, dtype):
    """A merge level whose interface matrix is tall enough for the unpivoted
    panel kernel (1792 unknowns, children with 448 Gauss nodes per side) on
    well-conditioned synthetic maps: the unpivoted factors pass the check
    and S, T and g_tilde equal the partially pivoted merge's (a single
    matrix: every panel unpivoted)."""
    from jaxhps_amd._stages import merge_factors_stable, merge_level
    from jaxhps_amd._tables import quad_merge_tables

    print(f"unpivoted tall merge {kind} {dtype}", flush=True)
    m = 448
    ne = 4 * m
    g = torch.Generator(device=gpu).manual_seed(3)
    real = torch.float64
    X = torch.randn(4, ne, ne, generator=g, device=gpu, dtype=real) / ne**0.5
    if kind == "iti":
        # contractive children: impedance maps of lossless problems are unitary
        Y = torch.randn(4, ne, ne, generator=g, device=gpu, dtype=real) / ne**0.5
        Tc = 0.5 * torch.complex(X, Y)
    else:
        Tc = 2.0 * torch.eye(ne, device=gpu, dtype=real) + 0.1 * X
    T = Tc.to(dtype).contiguous()
    h = torch.randn(4, ne, 2, generator=g, device=gpu, dtype=real).to(dtype).contiguous()
    tb = quad_merge_tables(m, kind)
    ref = merge_level(T, h, tb, pivot=True)
    out = merge_level(T, h, tb, pivot=False)
    assert bool(merge_factors_stable(out.F_lu))
    tol = 1e-11 if dtype in (torch.float64, torch.complex128) else 2e-4
    for a, b in ((out.S, ref.S), (out.T, ref.T), (out.gt, ref.gt), (out.h, ref.h)):
        err = float((a - b).abs().max() / b.abs().max())
        assert err < tol, err
