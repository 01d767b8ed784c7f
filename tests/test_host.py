"""Host-side pieces of jaxhps_amd on the CPU: spectral setup, point sets,
operators, API validation, merge index tables and the C ABI exports."""

import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

import cases2d as cs
from oracle import hps2d
from oracle import spectral as osp

ROOT = Path(__file__).resolve().parents[1]


def test_product_quadrature_textbook():
    from jaxhps_amd import quadrature as qd

    assert np.allclose(qd.chebyshev_points(3), [-1, 0, 1])
    D3 = np.array([[-1.5, 2.0, -0.5], [-0.5, 0.0, 0.5], [0.5, -2.0, 1.5]])
    assert np.allclose(qd.differentiation_matrix_1D(qd.chebyshev_points(3)), D3)
    assert np.allclose(qd.gauss_points(5), np.polynomial.legendre.leggauss(5)[0])
    assert np.allclose(qd.barycentric_lagrange_interpolation_matrix_1D(np.array([-1.0, 1.0]), np.array([0.0])), [[0.5, 0.5]])
    # Clenshaw-Curtis integrates polynomials of degree n-1 exactly
    for n in (5, 8, 12):
        x = qd.chebyshev_points(n)
        w = qd.chebyshev_weights(n, (-1.0, 1.0))
        for d in range(n):
            exact = (1 - (-1) ** (d + 1)) / (d + 1)
            assert abs(w @ x**d - exact) < 1e-12


@pytest.mark.parametrize("p,q", [(6, 4), (8, 6), (16, 14)])
def test_product_operators_match_oracle(p, q):
    from jaxhps_amd import _precompute_operators as po

    half = 0.37
    ops = osp.leaf_ops_2D(p, q, half, eta=1.5)
    dx, dy, dxx, dyy, dxy = po.precompute_diff_operators_2D(p, half)
    for a, b in [(dx, "D_x"), (dy, "D_y"), (dxx, "D_xx"), (dyy, "D_yy"), (dxy, "D_xy")]:
        assert np.allclose(a, ops[b], atol=1e-10 * np.abs(ops[b]).max())
    assert np.allclose(po.precompute_P_2D_DtN(p, q), ops["P"], atol=1e-12)
    assert np.allclose(po.precompute_Q_2D_DtN(p, q, dx, dy), ops["Q"], atol=1e-9)
    assert np.allclose(po.precompute_P_2D_ItI(p, q), ops["P_iti"], atol=1e-12)
    N = po.precompute_N_matrix_2D(dx, dy, p)
    Nt = po.precompute_N_tilde_matrix_2D(dx, dy, p)
    assert np.allclose(po.precompute_G_2D_ItI(Nt, 1.5), ops["G"], atol=1e-9)
    assert np.allclose(po.precompute_QH_2D_ItI(N, p, q, 1.5), ops["QH"], atol=1e-9)


def test_domain_points_match_oracle():
    import jaxhps_amd as jh

    for L in (0, 1, 2):
        dom = jh.Domain(p=6, q=4, root=jh.DiscretizationNode2D(*cs.BOUNDS), L=L)
        assert np.allclose(dom.interior_points, osp.leaf_points_2D(cs.BOUNDS, L, 6))
        assert np.allclose(dom.boundary_points, osp.boundary_gauss_points_2D(cs.BOUNDS, L, 4))
        assert dom.n_leaves == 4**L


def test_domain_3d_shapes():
    import jaxhps_amd as jh

    dom = jh.Domain(p=4, q=2, root=jh.DiscretizationNode3D(-1, 1, -1, 1, -1, 1), L=1)
    assert dom.interior_points.shape == (8, 64, 3)
    assert dom.boundary_points.shape == (6 * 4 * 4, 3)
    bp = dom.boundary_points
    assert np.allclose(bp[:16, 0], -1) and np.allclose(bp[16:32, 0], 1)
    assert np.allclose(bp[32:48, 1], -1) and np.allclose(bp[80:, 2], 1)


def test_pdeproblem_validation():
    import jaxhps_amd as jh

    dom = jh.Domain(p=6, q=4, root=jh.DiscretizationNode2D(-1, 1, -1, 1), L=1)
    ones = np.ones((4, 36))
    with pytest.raises(ValueError):
        jh.PDEProblem(dom, source=ones, D_xx_coefficients=ones, D_z_coefficients=ones)
    with pytest.raises(ValueError):
        jh.PDEProblem(dom, source=ones, D_xx_coefficients=ones, use_ItI=True)
    with pytest.raises(ValueError):
        jh.PDEProblem(dom, source=ones, D_xx_coefficients=np.ones((4, 35)))
    dom3 = jh.Domain(p=4, q=2, root=jh.DiscretizationNode3D(-1, 1, -1, 1, -1, 1), L=1)
    with pytest.raises(NotImplementedError):
        jh.PDEProblem(dom3, source=np.ones((8, 64)), D_xx_coefficients=np.ones((8, 64)), use_ItI=True, eta=1.0)
    with pytest.raises(ValueError):
        jh.PDEProblem(dom3, D_xx_coefficients=np.ones((8, 64)))
    pde = jh.PDEProblem(dom, source=ones, D_xx_coefficients=ones, D_yy_coefficients=ones)
    pde.S_lst = [1]
    pde.update_coefficients(I_coefficients=2 * ones)
    assert pde.S_lst == [] and pde.I_coefficients is not None


# ---------------------------------------------------------------------------
# merge tables: numpy emulation of the kernels' index semantics vs the oracle
# ---------------------------------------------------------------------------
def _gather(Tc, rows, cols):
    out = np.zeros((rows.shape[0], cols.shape[0]), dtype=Tc.dtype)
    for r in range(rows.shape[0]):
        for c in range(cols.shape[0]):
            acc = 0
            for a in (0, 2):
                if rows[r, a] < 0:
                    continue
                for b in (0, 2):
                    if cols[c, b] == rows[r, a]:
                        acc += Tc[rows[r, a], rows[r, a + 1], cols[c, b + 1]]
            out[r, c] = acc
    return out


def _vgather(hc, rows):
    out = np.zeros((rows.shape[0], hc.shape[-1]), dtype=hc.dtype)
    for r in range(rows.shape[0]):
        for a in (0, 2):
            if rows[r, a] >= 0:
                out[r] += hc[rows[r, a], rows[r, a + 1]]
    return out


def _emulate_merge(Tc, hc, tb):
    ni = tb.ni
    Y = -_gather(Tc, tb.int_rows, tb.ext)
    Yh = -_vgather(hc, tb.int_rows)
    if tb.kind == "dtn":
        D = _gather(Tc, tb.int_rows, tb.int_cols)
        X, Xh = np.linalg.solve(D, Y), np.linalg.solve(D, Yh)
    else:
        nF = ni // 2
        D12 = _gather(Tc, tb.int_rows[:nF], tb.int_cols[nF:])
        D21 = _gather(Tc, tb.int_rows[nF:], tb.int_cols[:nF])
        # the coupling rectangles hold every nonzero and cover each row once
        for blk, Dm in ((tb.d12_blk, D12), (tb.d21_blk, D21)):
            if len(blk):
                Ds, seen = np.zeros_like(Dm), np.zeros(nF, dtype=int)
                for r0, nr, k0, nk in blk:
                    Ds[r0:r0 + nr, k0:k0 + nk] = Dm[r0:r0 + nr, k0:k0 + nk]
                    seen[r0:r0 + nr] += 1
                assert np.array_equal(Ds, Dm) and (seen == 1).all()
                assert Ds.size > 2 * sum(int(nr) * int(nk) for _, nr, _, nk in blk) - 1  # half the block is zero
        W = np.eye(nF) - D12 @ D21
        out = []
        for y in (Y, Yh):
            xF = np.linalg.solve(W, y[:nF] - D12 @ y[nF:])
            out.append(np.concatenate([xF, y[nF:] - D21 @ xF]))
        X, Xh = out
    S = np.zeros_like(X)
    gt = np.zeros_like(Xh)
    S[tb.final_of_internal] = X
    gt[tb.final_of_internal] = Xh
    B = _gather(Tc, tb.ext, tb.cols_final)
    # the block-sparse product must equal the dense one
    Bs = np.zeros_like(B)
    for r0, nr, k0, nk in tb.upd:
        Bs[r0:r0 + nr, k0:k0 + nk] = B[r0:r0 + nr, k0:k0 + nk]
    assert np.array_equal(Bs, B)
    T = _gather(Tc, tb.ext, tb.ext) + B @ S
    h = _vgather(hc, tb.ext) + B @ gt
    return S, T, h, gt


@pytest.mark.parametrize("kind", ["dtn", "iti"])
@pytest.mark.parametrize("m", [1, 3, 64])
def test_merge_tables_emulation_matches_oracle(kind, m):
    from jaxhps_amd._tables import quad_merge_tables

    rng = np.random.default_rng(m)
    Tc = rng.standard_normal((4, 4 * m, 4 * m))
    if kind == "iti":
        Tc = 0.3 * (Tc + 1j * rng.standard_normal(Tc.shape))
    hc = rng.standard_normal((4, 4 * m, 2)) + 0j
    tb = quad_merge_tables(m, kind)
    S, T, h, gt = _emulate_merge(Tc, hc, tb)
    fn = hps2d.quad_merge_dtn if kind == "dtn" else hps2d.quad_merge_iti
    S0, T0, h0, g0 = fn(*Tc, *hc)
    for a, b in [(S, S0), (T, T0), (h, h0), (gt, g0)]:
        assert np.abs(a - b).max() < 1e-10 * max(1.0, np.abs(b).max())
    # down pass scatter
    g = rng.standard_normal((8 * m, 2))
    gint = S @ g + gt
    out = np.zeros((4, 4 * m, 2), dtype=complex)
    for slot, src in enumerate(tb.child_src):
        c, l = divmod(slot, 4 * m)
        out[c, l] = g[src] if src >= 0 else gint[-src - 1]
    dfn = hps2d.propagate_down_dtn if kind == "dtn" else hps2d.propagate_down_iti
    assert np.abs(out - dfn(S0, g, g0)).max() < 1e-10


# ---------------------------------------------------------------------------
# C ABI
# ---------------------------------------------------------------------------
def test_library_exports_every_declared_symbol():
    from jaxhps_amd import _lib

    header = (ROOT / "include" / "hps_mi355x.h").read_text()
    names = set(re.findall(r"^\s*(?:int|size_t|const char\*)\s+(hps_\w+)\s*\(", header, re.M))
    assert len(names) >= 15
    lib = _lib.load_library()
    for nm in names:
        assert hasattr(lib, nm), nm
        assert nm in _lib.SIGNATURES, f"{nm} has no ctypes signature"
    assert lib.hps_version() == 1


def test_ctypes_signatures_match_header_arity():
    """Every ctypes signature has as many arguments as the header declaration."""
    from jaxhps_amd import _lib

    header = (ROOT / "include" / "hps_mi355x.h").read_text()
    header = re.sub(r"/\*.*?\*/", "", header, flags=re.S)
    decls = re.findall(r"^\s*(?:int|size_t|const char\*)\s+(hps_\w+)\s*\(([^)]*)\)\s*;", header, re.M)
    assert len(decls) >= 15
    for name, params in decls:
        params = params.strip()
        n = 0 if params in ("", "void") else len(params.split(","))
        assert len(_lib.SIGNATURES[name][1]) == n, (name, n, len(_lib.SIGNATURES[name][1]))


def test_library_argument_errors_without_gpu():
    """Argument checks run before any device work, so they are testable on the CPU."""
    from jaxhps_amd import _lib

    lib = _lib.load_library()
    h = ctypes.c_void_p()
    assert lib.hps_create(None, ctypes.byref(h)) == 0
    rc = lib.hps_gemm_batched(h, 0, 4, 4, 4, 1.0, None, 4, 0, None, 4, 0, 0.0, None, 4, 0, None, 4, 0, 1)
    assert rc == -1 and b"null" in lib.hps_last_error(h)
    rc = lib.hps_getrf_batched(h, 0, 20000, None, 20000, 0, None, None, None, 1, None, 0)
    assert rc == -1
    rc = lib.hps_gemm_batched(h, 9, 4, 4, 4, 1.0, ctypes.c_void_p(8), 4, 0, ctypes.c_void_p(8), 4, 0, 0.0, None, 4,
                              0, ctypes.c_void_p(8), 4, 0, 1)
    assert rc == -1 and b"dtype" in lib.hps_last_error(h)
    assert lib.hps_destroy(h) == 0


def test_block_inverse_sizes_and_argument_errors():
    """hps_block_inverses: eligible sizes (n > 256, a multiple of 32) hold two
    256 x 256 inverses per diagonal block (the last one padded), others none;
    bad sizes and pointers are refused before any device work."""
    from jaxhps_amd import _lib

    lib = _lib.load_library()
    assert lib.hps_block_inverses_elems(256) == 0
    assert lib.hps_block_inverses_elems(300) == 0
    assert lib.hps_block_inverses_elems(288) == 2 * 2 * 256 * 256
    assert lib.hps_block_inverses_elems(3584) == 2 * 14 * 256 * 256
    ws = lib.hps_block_inverses_workspace(_lib.HPS_C64, 3584, 2)
    assert ws >= 2 * 14 * (256 * 256 + 256 * 32) * 8
    h = ctypes.c_void_p()
    assert lib.hps_create(None, ctypes.byref(h)) == 0
    p = ctypes.c_void_p(8)
    rc = lib.hps_block_inverses(h, _lib.HPS_C64, 256, 1, p, p, p, p, 1 << 30)
    assert rc == -1 and b"n > 256" in lib.hps_last_error(h)
    rc = lib.hps_block_inverses(h, _lib.HPS_C64, 288, 0, p, p, p, p, 1 << 30)
    assert rc == -1 and b"count" in lib.hps_last_error(h)
    rc = lib.hps_block_inverses(h, _lib.HPS_C64, 288, 1, None, p, p, p, 1 << 30)
    assert rc == -1 and b"null" in lib.hps_last_error(h)
    rc = lib.hps_block_inverses(h, _lib.HPS_C64, 288, 1, p, p, p, p, 16)
    assert rc != 0 and b"workspace" in lib.hps_last_error(h)
    assert lib.hps_destroy(h) == 0


def test_product_fails_loudly_without_gpu():
    import torch

    import jaxhps_amd as jh

    if torch.cuda.is_available():
        pytest.skip("GPU present")
    dom = jh.Domain(p=6, q=4, root=jh.DiscretizationNode2D(-1, 1, -1, 1), L=1)
    ones = np.ones((4, 36))
    pde = jh.PDEProblem(dom, source=ones, D_xx_coefficients=ones, D_yy_coefficients=ones)
    with pytest.raises(jh.HPSLibraryError):
        jh.build_solver(pde)


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU; --gpus above 1 exits with a message."""
    import subprocess
    import sys

    root = Path(__file__).resolve().parents[1]
    r = subprocess.run([sys.executable, str(root / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_scattering_and_interp_fail_loudly_without_gpu():
    """The exterior coupling and interpolation have no CPU path either."""
    import torch

    if torch.cuda.is_available():
        pytest.skip("needs a host without a GPU")
    import jaxhps_amd as jh
    from jaxhps_amd import HPSLibraryError
    from jaxhps_amd.scattering import get_DtN_from_ItI

    with pytest.raises((HPSLibraryError, RuntimeError)):
        get_DtN_from_ItI(np.eye(8, dtype=complex) * 0.5, 3.0)
    dom = jh.Domain(p=4, q=2, root=jh.DiscretizationNode2D(0.0, 1.0, 0.0, 1.0), L=1)
    with pytest.raises((HPSLibraryError, RuntimeError)):
        dom.interp_from_interior_points(np.zeros(dom.interior_points.shape[:2]), np.linspace(0, 1, 3),
                                        np.linspace(0, 1, 3))


def test_gemm_wrapper_rejects_mismatched_operands():
    """The batched GEMM indexes every operand by the common batch: shapes that
    would make it read past an operand's end are refused before any launch."""
    import torch

    from jaxhps_amd import _lib

    A = torch.zeros(3, 4, 5, dtype=torch.float64)
    with pytest.raises(ValueError):
        _lib.gemm(A, torch.zeros(5, 5, 6, dtype=torch.float64))  # batch 3 vs 5
    with pytest.raises(ValueError):
        _lib.gemm(A, torch.zeros(3, 4, 6, dtype=torch.float64))  # inner 5 vs 4
    with pytest.raises(ValueError):
        _lib.gemm(A, torch.zeros(3, 5, 6, dtype=torch.float64), C=torch.zeros(2, 4, 6, dtype=torch.float64))
    with pytest.raises(ValueError):
        _lib.gemm(A, torch.zeros(3, 5, 6, dtype=torch.float32))


def test_unpivoted_factor_check():
    """merge_factors_stable accepts natural-order LU factors with small
    multipliers and rejects a zero pivot, a large multiplier or a
    non-finite entry (CPU tensors; the check is plain torch)."""
    import torch

    from jaxhps_amd._stages import merge_factors_stable

    rng = np.random.default_rng(0)
    A = np.eye(6) + 0.1 * rng.standard_normal((6, 6))
    Lu = torch.as_tensor(np.stack([A, A]))
    assert bool(merge_factors_stable(Lu))
    bad = Lu.clone()
    bad[1, 3, 3] = 0.0
    assert not bool(merge_factors_stable(bad))
    bad = Lu.clone()
    bad[0, 4, 1] = 100.0
    assert not bool(merge_factors_stable(bad))
    bad = Lu.clone().to(torch.complex128)
    bad[0, 2, 5] = complex("nan")
    assert not bool(merge_factors_stable(bad))
