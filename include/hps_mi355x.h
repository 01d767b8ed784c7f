/*
 * hps_mi355x.h -- C ABI of libhps_mi355x.so, the MI355X (gfx950) kernels
 * behind the uniform-tree HPS build_solver()/solve() hot path.
 *
 * The reference (jaxhps) is pure Python/JAX and has no FFI; each entry point
 * below replaces the batched XLA computation of one reference stage function,
 * cited per declaration (paths relative to the reference's src/jaxhps/).
 *
 * Conventions
 *   - All array arguments are DEVICE pointers (HBM), row-major, leaf/merge
 *     batch axis first, exactly the layout of the reference's arrays.
 *   - dtype: HPS_F64, HPS_F32, HPS_C128, HPS_C64 (complex = interleaved re,im).
 *     Every floating-point array of a call has that dtype.
 *   - Calls are asynchronous on the handle's stream; the caller synchronises.
 *     A handle is bound to one stream and is not thread-safe.
 *   - Workspace is caller-provided device memory; *_workspace() returns the
 *     size in bytes.  The library never allocates or frees device memory.
 *   - Return value: HPS_OK, or a negative error (hps_last_error() explains).
 *     Numerical breakdown (an exactly zero pivot column) is reported per batch
 *     element in the device int array `info` (0 = ok, j+1 = column j), the
 *     LAPACK getrf convention.
 */
#ifndef HPS_MI355X_H
#define HPS_MI355X_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HPS_ABI_VERSION 1

#define HPS_OK 0
#define HPS_ERR_ARG (-1)
#define HPS_ERR_HIP (-2)
#define HPS_ERR_WORKSPACE (-3)

#define HPS_F64 0
#define HPS_F32 1
#define HPS_C128 2
#define HPS_C64 3

#define HPS_KIND_DTN 0
#define HPS_KIND_ITI 1

typedef struct hps_handle_s* hps_handle_t;

/* Handle bound to a hipStream_t (NULL = default stream). */
int hps_create(void* stream, hps_handle_t* out);
int hps_destroy(hps_handle_t h);
int hps_set_stream(hps_handle_t h, void* stream);
const char* hps_last_error(hps_handle_t h);
int hps_version(void);

/* ---- batched dense primitives --------------------------------------------
 * D[b] = alpha*A[b]@B[b] + beta*C[b]; batch stride 0 broadcasts.  Replaces
 * the vmapped matmuls/einsums of the reference, e.g. T = Q @ Y
 * (local_solve/_uniform_2D_DtN.py:267) and the leaf apply einsum
 * (down_pass/_uniform_2D_DtN.py:114-120). */
int hps_gemm_batched(hps_handle_t h, int dtype, int M, int N, int K, double alpha, const void* A, long lda,
                     long sA, const void* B, long ldb, long sB, double beta, const void* C, long ldc, long sC,
                     void* D, long ldd, long sD, int batch);

/* In-place LU with partial pivoting, P A = L U; perm[b*n+i] = source row of
 * row i; dinv[b] (hps_dinv_size(n) elements per matrix) = packed inverses of
 * the 32x32 diagonal blocks (strict lower: inv(L_kk), upper: inv(U_kk)),
 * which turn every triangular solve into parallel block products.
 * n <= 12288.  Replaces the factorisation inside jnp.linalg.inv
 * (local_solve/_uniform_2D_DtN.py:257, merge/_schur_complement.py:107,146). */
size_t hps_getrf_workspace(int n, int batch);
size_t hps_dinv_size(int n);
int hps_getrf_batched(hps_handle_t h, int dtype, int n, void* A, long lda, long sA, int* perm, void* dinv, int* info,
                      int batch, void* ws, size_t ws_bytes);

/* X = A^{-1} R from hps_getrf_batched's factors (out of place, X != R).
 * Replaces the products with inv(A) (local_solve/_uniform_2D_DtN.py:260-270). */
int hps_getrs_batched(hps_handle_t h, int dtype, int n, int nrhs, const void* LU, long lda, long sA, const int* perm,
                      const void* dinv, const void* R, long ldr, long sR, void* X, long ldx, long sX, int batch);

/* ---- leaf stage -----------------------------------------------------------
 * Batched leaf solve over nl leaves of n = p^d collocation points, nb
 * Chebyshev boundary points, nq Gauss boundary points and nsrc sources.
 *   coeffs   (ncoef, nl, n)   coefficient fields that are present
 *   coef_ops host int[ncoef]: index of each coefficient's operator in
 *            diffops (nops, n, n), or -1 for the identity (I_coefficients)
 *   P (nb, nq), Q (nq, n) [DtN: Q, ItI: QH], G (nb, n) [ItI only]
 *   source   (nl, n, nsrc) or NULL (zero source)
 * Outputs Y (nl, n, nq), T (nl, nq, nq), v (nl, n, nsrc), h (nl, nq, nsrc),
 *   lu:   DtN (nl, n-nb, n) = [A_ie | LU(A_ii)];  ItI (nl, n, n) = LU(B)
 *   perm: DtN (nl, n-nb);  ItI (nl, n);  dinv: (nl, hps_dinv_size(rows));
 *   info (nl).
 * DtN replaces local_solve_stage_uniform_2D_DtN + get_DtN
 *   (local_solve/_uniform_2D_DtN.py:9-102,182-273) and the 3D version
 *   (local_solve/_uniform_3D_DtN.py:13-106);
 * ItI replaces local_solve_stage_uniform_2D_ItI + get_ItI
 *   (local_solve/_uniform_2D_ItI.py:10-193).
 * With nsrc = 0 it is the no-source leaf stage
 *   (local_solve/_nosource_uniform_2D_{DtN,ItI}.py); Phi is kept as lu/perm. */
size_t hps_leaf_workspace(int dtype, int kind, int nl, int n, int nb, int nq, int nsrc);
int hps_leaf_solve(hps_handle_t h, int dtype, int kind, int nl, int n, int nb, int nq, int nsrc, int ncoef,
                   const int* coef_ops, const void* coeffs, const void* diffops, const void* P, const void* Q,
                   const void* G, const void* source, void* Y, void* T, void* v, void* hvec, void* lu, int* perm,
                   void* dinv, int* info, void* ws, size_t ws_bytes);

/* ---- merge stage (one tree level) ------------------------------------------
 * nm merges of nc children each (4 in 2D, 8 in 3D); children's maps
 * T_in (nm*nc, nec, nec) and data h_in (nm*nc, nec, nsrc).
 * Index tables (device int32, built once per level on the host):
 *   int_rows (ni, 4): for interface equation r, up to two (child, local) pairs
 *                     whose outgoing data it sums (child -1 = unused)
 *   int_cols (ni, 4): for interface unknown j, the (child, local) slots it feeds
 *   ext      (ne, 4): (child, local, -1, 0) per merged exterior node, in the
 *                     merged node's final order (the reference's roll folded in)
 *   cols_final (ni, 4): int_cols in the final unknown order (B's columns);
 *                     NULL = int_cols
 *   final_of_internal (ni): final row of each internal unknown (ItI only)
 * The interface system D g_int = -(C g_ext + h_int), D = gather(rows, cols):
 *   DtN: D is dense; F = LU(D), F (nm, ni, ni).
 *   ItI: unknowns in internal order [F-half; G-half] give
 *        D = I + [[0, D12], [D21, 0]]; F = LU(W), W = I - D12 D21
 *        (nm, ni/2, ni/2), and D^{-1} is applied through the Schur complement
 *        (merge/_schur_complement.py:79-114).  D12/D21_out (nm, ni/2, ni/2)
 *        optionally keep the blocks.
 * Outputs (final order):  S_out (nm, ni, ne) = -D^{-1} C,
 *   gt_out (nm, ni, nsrc) = -D^{-1} h_int,  T_out (nm, ne, ne) = A + B S,
 *   h_out (nm, ne, nsrc) = h_ext + B gt,  perm/F_dinv/info of F,
 *   B_out (nm, ne, ni) optional copy of B.  T_out/h_out may be NULL.
 *   F_inv_out (nm * hps_block_inverses_elems(n), n = ni (DtN) or ni/2 (ItI))
 *   or NULL: for complex64 / float32 levels of fewer than 128 matrices with
 *   hps_block_inverses_elems(n) > 0, the inverses of F's diagonal blocks (as
 *   hps_block_inverses forms them), which the level's few-column source
 *   solve uses too; otherwise not written.
 * upd (host int[4*n_upd]): the nonzero blocks (r0, nr, k0, nk) of B; the
 *   products B S and B gt use only these (n_upd = 0: B dense).
 * d12_blk / d21_blk (host int[4*n], ItI only): the nonzero rectangles of D12
 *   and D21 (an F unknown's equation sees only the G unknowns of its
 *   partner child and vice versa), each row in exactly one rectangle; the
 *   products D12 D21, D12 Y_G and D21 X_F use only these (n = 0: dense).
 * Replaces _uniform_quad_merge_DtN / _ItI with assemble_merge_outputs_*
 *   (merge/_uniform_2D_DtN.py:207-348, merge/_uniform_2D_ItI.py:183-375,
 *   merge/_schur_complement.py:7-237) and _uniform_oct_merge_DtN
 *   (merge/_uniform_3D_DtN.py:128-187, merge/_schur_complement.py:294-774). */
/* Pivoting of hps_merge_level's factorisation of D (DtN) or W (ItI) on this
 * handle: 1 (default) = partial pivoting; 0 = every panel is factored
 * without row exchanges.  The merge matrices of the uniform tree eliminate
 * stably in their natural order (pivot growth ~1, multipliers below 4 on the
 * Helmholtz and variable-coefficient workloads), while pivoted panels are
 * latency-bound.  With 0, the register panels (up to 1024 rows, 512 for
 * complex128, in levels of more than 16 matrices) check each matrix's
 * largest multiplier and set info = -1 when it exceeds 64 or is not finite;
 * the other panels are not checked, and a caller that turns pivoting off
 * checks those factors (largest multiplier, finite, nonzero diagonal) and
 * refactors with pivoting when either check fails.  The reference factors
 * with LAPACK's pivoted LU inside jnp.linalg.inv
 * (merge/_schur_complement.py:107,146). */
int hps_set_merge_pivoting(hps_handle_t h, int partial_pivoting);
/* Pivoting of hps_leaf_solve's factorisation on this handle: 1 (default) =
 * partial pivoting; 0 = natural order, every panel checked in the kernel:
 * a leaf whose largest multiplier exceeds 64 (or is not finite) gets
 * info[l] = -1, and the caller reruns the leaf stage with pivoting.  The
 * row-scaled ItI leaf matrices and the DtN interior blocks eliminate with
 * multipliers below 2 on the benchmark workloads.  The reference factors
 * with LAPACK's pivoted LU inside jnp.linalg.inv
 * (local_solve/_uniform_2D_ItI.py:167, local_solve/_uniform_2D_DtN.py:257). */
int hps_set_leaf_pivoting(hps_handle_t h, int partial_pivoting);
size_t hps_merge_workspace(int dtype, int kind, int nm, int ni, int ne, int nsrc);
int hps_merge_level(hps_handle_t h, int dtype, int kind, int nm, int nc, int nec, int ni, int ne, int nsrc,
                    const void* T_in, const void* h_in, const int* int_rows, const int* int_cols, const int* ext,
                    const int* cols_final, const int* final_of_internal, int n_upd, const int* upd, int n_d12,
                    const int* d12_blk, int n_d21, const int* d21_blk, void* S_out, void* T_out, void* h_out,
                    void* gt_out, void* F_lu, int* perm, void* F_dinv, int* info, void* D12_out, void* D21_out,
                    void* B_out, void* F_inv_out, void* ws, size_t ws_bytes);

/* ---- up pass (one tree level) --------------------------------------------
 * New outgoing data through the factors a merge level kept (hps_merge_level
 * with nsrc = 0 and F_lu/perm/F_dinv, D12/D21 (ItI) and B_out requested):
 *   h_int = gather(h_in, int_rows) + inj,  gt_out = -D^{-1} h_int (final
 *   order),  h_out = h_ext + B gt_out (h_out may be NULL at the root).
 * inj (nm, ni, nsrc), internal equation order, or NULL: extra interface data
 * (the interface residual of iterative refinement).  F_inv: the block
 * inverses hps_block_inverses formed from F_lu, or NULL; with them (and
 * nsrc <= 16) the solve with D (W) runs through them.
 * Replaces up_pass_uniform_2D_ItI / _DtN with assemble_boundary_data
 * (up_pass/_uniform_2D_ItI.py:8-219, up_pass/_uniform_2D_DtN.py:8-173). */
size_t hps_up_workspace(int dtype, int kind, int nm, int ni, int nsrc);
int hps_up_level(hps_handle_t h, int dtype, int kind, int nm, int nc, int nec, int ni, int ne, int nsrc,
                 const void* h_in, const void* inj, const int* int_rows, const int* ext, const int* final_of_internal,
                 int n_upd, const int* upd, const void* F_lu, const int* perm, const void* F_dinv,
                 const void* F_inv, const void* D12, const void* D21, const void* B, void* gt_out, void* h_out,
                 void* ws, size_t ws_bytes);

/* Inverses of the 256 x 256 diagonal blocks of L (unit lower) and U of the
 * nm factored n x n merge matrices F_lu / F_dinv of hps_merge_level, for
 * hps_up_level's few-column solves: each 256-row block step becomes one
 * product with a block inverse instead of eight dependent 32-row steps.
 * inv holds nm * hps_block_inverses_elems(n) elements (0 when n is not
 * eligible: n > 256, a multiple of 32).  No counterpart in the reference,
 * whose up pass applies the explicit inverse D^{-1} it stores
 * (up_pass/_uniform_2D_ItI.py:180-219); the rounding of the inverted blocks
 * is that of the low-precision factors, corrected by the refinement. */
size_t hps_block_inverses_elems(int n);
size_t hps_block_inverses_workspace(int dtype, int n, int nm);
int hps_block_inverses(hps_handle_t h, int dtype, int n, int nm, const void* F_lu, const void* F_dinv, void* inv,
                       void* ws, size_t ws_bytes);

/* ---- down pass (one tree level) -------------------------------------------
 * g_int = S g_in + gt; children's boundary data g_out (nm*nc, nec, nsrc) from
 * child_src (nc*nec ints): >= 0 parent exterior index, < 0 interface index
 * -(j+1).  Replaces _propagate_down_2D_DtN / _propogate_down_2D_ItI /
 * _propogate_down_oct_DtN (down_pass/_uniform_2D_DtN.py:126-189,
 * down_pass/_uniform_2D_ItI.py:122-192, down_pass/_uniform_3D_DtN.py:117-246). */
size_t hps_down_workspace(int dtype, int nm, int ni, int nsrc);
int hps_down_level(hps_handle_t h, int dtype, int nm, int nc, int nec, int ni, int ne, int nsrc, const void* S,
                   const void* gt, const void* g_in, const int* child_src, void* g_out, void* ws, size_t ws_bytes);

/* hps_down_level with no workspace argument and no allocation: g_int lives
 * in a library-owned static device array of HPS_STATIC_SCRATCH_BYTES per GPU
 * (hps_static_scratch_bytes()), held until the process ends.  No counterpart
 * in the reference: it is for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops re-solving for new
 * boundary data).  Sizes with nm * ni * nsrc * sizeof(element) above the
 * array return HPS_ERR_WORKSPACE.  Every call on one GPU shares that GPU's
 * array: calls that overlap on one GPU (several streams or threads) race.
 * The handle must have been created on the GPU it runs on. */
#define HPS_STATIC_SCRATCH_BYTES (32u << 20)
size_t hps_static_scratch_bytes(void);
int hps_down_level_static(hps_handle_t h, int dtype, int nm, int nc, int nec, int ni, int ne, int nsrc,
                          const void* S, const void* gt, const void* g_in, const int* child_src, void* g_out);

/* u = Y g + v per leaf (v may be NULL).  Replaces the leaf einsum of the down
 * pass (down_pass/_uniform_2D_DtN.py:110-120). */
int hps_leaf_apply(hps_handle_t h, int dtype, int nl, int n, int nq, int nsrc, const void* Y, const void* g,
                   const void* v, void* u);

/* ---- iterative refinement --------------------------------------------------
 * The reference has no mixed-precision path; these entry points serve the
 * FP32/complex64 build with FP64 refinement.  Each evaluates, in FP64, the
 * residual of the global system the uniform build+solve solves exactly:
 * the leaf rows of get_DtN / get_ItI (local_solve/_uniform_2D_DtN.py:228-280,
 * local_solve/_uniform_2D_ItI.py:120-193) and the interface conditions the
 * merges eliminate (merge/_schur_complement.py:183-237 for DtN flux balance,
 * merge/_uniform_2D_ItI.py:183-375 for ItI impedance exchange).
 *
 * R (nl, nrow_R, nsrc): [op_0 u | ... | op_{nops-1} u | extra], one GEMM of
 * the stacked operators with the iterate u (nl, n, nsrc); extra = [G u | QH u]
 * (ItI, nrow_R = nops*n + nb + nq) or [Q u] (DtN, nrow_R = nops*n + nq).
 * coeffs (ncoef, nl, n) with coef_ops[k] in [-1, nops) (-1 = identity).
 * nbr (nl, nq): neighbour slot l'*nq + j' (>= 0) or -(root index + 1).
 * out_dtype is dtype or its FP32 partner (the correction's precision);
 * part (nl) receives per-leaf sums of |residual|^2. */
int hps_residual_iti(hps_handle_t h, int dtype, int out_dtype, int nl, int n, int nb, int nq, int nsrc, int nops,
                     int ncoef, const int* coef_ops, const void* coeffs, const void* R, int nrow_R, const void* u,
                     const void* f, const void* P, const void* g_root, const int* nbr, void* res, double* part);
/* DtN: u's boundary rows must equal P g_leaf (g_leaf (nl, nq, nsrc), the
 * leaves' Gauss boundary data).  res (nl, n, nsrc): interior rows f - A u,
 * boundary rows 0; hinj (nl, nq, nsrc): interface flux mismatch on one slot
 * of each interior Gauss node; groot_res (nroot, nsrc): g_root - g_leaf. */
int hps_residual_dtn(hps_handle_t h, int dtype, int out_dtype, int nl, int n, int nb, int nq, int nsrc, int nops,
                     int ncoef, const int* coef_ops, const void* coeffs, const void* R, int nrow_R, const void* u,
                     const void* f, const void* g_leaf, const void* g_root, const int* nbr, void* res, void* hinj,
                     void* groot_res, double* part);
/* x += dx over n elements (dx in dtype or its FP32 partner); nparts blocks,
 * part (2*nparts): per-block sums of |dx|^2 and |x|^2. */
int hps_refine_update(hps_handle_t h, int dtype, int delta_dtype, long n, void* x, const void* dx, int nparts,
                      double* part);
/* dst = src converted elementwise (complex -> real keeps the real part). */
int hps_convert(hps_handle_t h, int src_dtype, int dst_dtype, long n, const void* src, void* dst);

/* ---- adjoint (reverse mode) through the stored factors --------------------
 * The reference differentiates build+solve with JAX reverse mode
 * (examples/check_autodiff_vJp.py:154-243, inverse_wave_scattering.py:209-292
 * use it); these entry points are the hand-written adjoint of the uniform
 * solve: the adjoint of each down-pass level, of each up-pass (merge) level
 * and of the leaf solves, so that lambda = K^{-H} r for the global system K
 * the build factored, and the gradient of the coefficient fields from lambda.
 *
 * D = alpha * A^H @ B + beta * C with A stored K x M (lda). */
int hps_gemm_batched_conjt(hps_handle_t h, int dtype, int M, int N, int K, double alpha, const void* A, long lda,
                           long sA, const void* B, long ldb, long sB, double beta, const void* C, long ldc, long sC,
                           void* D, long ldd, long sD, int batch);
/* X = A^{-H} R from hps_getrf_batched's factors (out of place). */
size_t hps_getrs_conjt_workspace(int dtype, int n, int nrhs, int batch);
int hps_getrs_batched_conjt(hps_handle_t h, int dtype, int n, int nrhs, const void* LU, long lda, long sA,
                            const int* perm, const void* dinv, const void* R, long ldr, long sR, void* X, long ldx,
                            long sX, int batch, void* ws, size_t ws_bytes);
/* Adjoint of hps_down_level: children's data gbar_in (nm*nc, nec, nsrc) ->
 * gtbar_out (nm, ni, nsrc) = sum over each interface unknown's slots
 * (cols_final table), gbar_ext_out (nm, ne, nsrc) = exterior slots + S^H gtbar. */
int hps_adjoint_down_level(hps_handle_t h, int dtype, int nm, int nc, int nec, int ni, int ne, int nsrc, const void* S,
                           const void* gbar_in, const int* ext, const int* cols_final, void* gbar_ext_out,
                           void* gtbar_out);
/* Adjoint of hps_up_level: y = gtbar + B^H hbar_out (hbar_out NULL at the
 * root), hbar_int = -D^{-H} y (internal order via final_of_internal; ItI in
 * the Schur form of W = I - D12 D21), children's slots from adj_src
 * (>= 0 exterior index into hbar_out, < 0 internal row -(r+1)). */
size_t hps_adjoint_up_workspace(int dtype, int kind, int nm, int ni, int nsrc);
int hps_adjoint_up_level(hps_handle_t h, int dtype, int kind, int nm, int nc, int nec, int ni, int ne, int nsrc,
                         const void* gtbar, const void* hbar_out, const void* B, const void* F_lu, const int* perm,
                         const void* F_dinv, const void* D12, const void* D21, const int* final_of_internal,
                         const int* adj_src, void* hbar_children, void* ws, size_t ws_bytes);
/* grad (ncoef, nl, n) float64: -Re sum_s conj(lam) * (op_k u) on rows >= row0
 * (the rows the coefficients enter), 0 below; R as for the residuals. */
int hps_coef_gradient(hps_handle_t h, int dtype, int nl, int n, int row0, int nsrc, int nops, int ncoef,
                      const int* coef_ops, const void* R, int nrow_R, const void* u, const void* lam, double* grad);


/* ---- adaptive (non-uniform) trees ------------------------------------------
 * X[b*stride + i] *= scale[b] (i < n), scale a device double array: the
 * per-leaf side-length scaling of coefficient fields and of DtN maps and
 * fluxes that replaces rebuilding the operators for every leaf size
 * (local_solve/_adaptive_2D_DtN.py:130-172). */
int hps_scale_batched(hps_handle_t h, int dtype, int batch, long n, long stride, const double* scale, void* X);

/* count block copies in one launch: desc (device, count x 6 int64) =
 * (src byte address, dst byte address, rows, cols, lds, ldd) per block, in
 * elements of dtype; max_elems >= rows * cols of every block (sizes the
 * grid).  Packs the children of a group of adaptive merges, and their
 * boundary data, without a launch per node (merge/_adaptive_2D_DtN.py:
 * 110-206, down_pass/_adaptive_2D_DtN.py:44-98). */
int hps_pack_blocks(hps_handle_t h, int dtype, int count, long max_elems, const long long* desc);

/* ---- exterior-scattering coupling of the top-level operator ---------------
 * Dense complex128 work on the n root-boundary Gauss points (S, E, N, W sides
 * of n/4 points each).  Device pointers; info is one device int (0, or j+1
 * for the first exactly singular pivot column of the dense LU).
 *
 * u_in[i, s] = exp(i k <pts_i, (cos dirs_s, sin dirs_s)>), (npts, nsrc); with
 * uin_dn, the outward normal derivative on boundary points in sides of
 * n_per_side points (the last side takes the remainder).  Replaces get_uin
 * and get_uin_and_normals (examples/wave_scattering_utils.py:139-197). */
int hps_plane_wave(hps_handle_t h, long npts, int nsrc, const double* pts, const double* dirs, double k,
                   long n_per_side, void* uin, void* uin_dn);

/* T = -i eta (R - I)^{-1} (R + I) for the n x n top-level ItI map R (T may
 * alias R).  Replaces get_DtN_from_ItI (examples/wave_scattering_utils.py:
 * 31-49). */
size_t hps_iti_to_dtn_workspace(int n);
int hps_iti_to_dtn(hps_handle_t h, int dtype, int n, const void* R, long ldr, double eta, void* T, long ldt,
                   int* info, void* ws, size_t ws_bytes);

/* A = I/2 - D + S T (n x n) and b = S (du_in/dn - T u_in) (n x nsrc) for the
 * boundary-integral equation of the exterior problem.  Replaces
 * setup_scattering_lin_system (examples/wave_scattering_utils.py:96-136). */
size_t hps_scattering_workspace(int n, int nsrc);
int hps_scattering_system(hps_handle_t h, int dtype, int n, int nsrc, const void* S, const void* D, const void* T,
                          const double* bdry_pts, const double* source_dirs, double k, void* A, void* b, void* ws,
                          size_t ws_bytes);

/* imp (n x nsrc) = du_scat/dn + i eta u_scat with u_scat = A^{-1} b and
 * du_scat/dn = T (u_scat + u_in) - du_in/dn: the incoming impedance data of
 * the scattered field for the down pass.  Workspace as
 * hps_scattering_workspace.  Replaces get_scattering_uscat_impedance
 * (examples/wave_scattering_utils.py:200-244). */
int hps_scattering_impedance(hps_handle_t h, int dtype, int n, int nsrc, const void* S, const void* D, const void* T,
                             const double* bdry_pts, const double* source_dirs, double k, double eta, void* imp,
                             int* info, void* ws, size_t ws_bytes);

/* Interpolation from the HPS leaf grids: F (n_leaves, p^dim, nsrc) in leaf
 * point order, boxes (n_leaves, 2 dim) = [lo, hi] per coordinate, cheb (p)
 * the Chebyshev nodes on [-1, 1] ascending, flip_y (2D) runs the leaf's y
 * nodes top down, grid_to_leaf (p^dim) maps the tensor-grid position (x
 * slowest) to the leaf's point order; barycentric weights with exact
 * coincidences replaced by EPS; FP64 / complex128.  Evaluated on the points
 * of meshgrid(xs, ys[, zs]) in its default
 * ("xy") order, flat index (iy * nx + ix) * nz + iz (nz = 1 in 2D), with
 * each point's leaf found on the device for a uniform tree: root (2 dim) =
 * [lo, hi] per coordinate, n_side leaves per axis, leaf_of_cell
 * (n_side^dim, x slowest) the leaf of each cell; a point on shared edges
 * takes the first leaf in leaf order, one outside every leaf takes leaf 0.
 * out (n_points, nsrc).  Replaces interp_from_hps_2D / _3D
 * (_interpolation_methods.py:24-219) for uniform trees. */
int hps_interp_grid(hps_handle_t h, int dtype, int nx, int ny, int nz, int dim, int p, int nsrc, const double* xs,
                    const double* ys, const double* zs, int n_side, const double* root, const int* leaf_of_cell,
                    const double* boxes, const double* cheb, int flip_y, const int* grid_to_leaf, const void* F,
                    void* out);

/* ---- learned leaf-ItI surrogate (no counterpart in the reference) ---------
 * A small MLP maps a leaf's coefficient samples to a rank-r complex
 * correction U V^H of the q = 0 leaf ItI map.  FP32 tensors in memory, BF16
 * MFMA products with FP32 accumulation.
 * C = op(A) op(B) [+ beta C] [+ bias[n]] [tanh]; op(A) = A^T (A stored K x M)
 * when transA, op(B) = B^T (B stored N x K) when transB. */
int hps_mlp_gemm_bf16(hps_handle_t h, int transA, int transB, int M, int N, int K, const float* A, int lda,
                      const float* B, int ldb, float beta, float* C, int ldc, const float* bias, int act);
/* dH *= 1 - H^2 (tanh backward, in place). */
int hps_mlp_tanh_backward(hps_handle_t h, long n, const float* H, float* dH);
/* db[n] = sum_m dZ[m][n]. */
int hps_mlp_bias_grad(hps_handle_t h, int M, int N, const float* dZ, float* db);
/* Adam step (bias-corrected, step >= 1) on n parameters. */
int hps_mlp_adam(hps_handle_t h, long n, float* w, const float* g, float* m1, float* m2, float lr, float b1, float b2,
                 float eps, int step);
/* Y (nsamp, 4 nq r) = [U re | U im | V re | V im]; dT (nsamp, nq, nq)
 * complex64: loss[s] = scale |U V^H - dT|_F^2, dY = its gradient. */
int hps_mlp_lowrank_loss(hps_handle_t h, int nsamp, int nq, int r, const float* Y, const float* dT, float scale,
                         float* dY, float* loss);
/* T[s] = T0 + out_scale U V^H, complex128 (nsamp, nq, nq). */
int hps_mlp_lowrank_apply(hps_handle_t h, int nsamp, int nq, int r, const float* Y, const double* T0,
                          double out_scale, double* T);

/* ---- optional kernel timing ----------------------------------------------
 * When enabled, every kernel launch of the library is bracketed by HIP events
 * on its stream and attributed to a kernel class ("gemm_mfma<c128>", ...)
 * with the launch's algorithmic flop and byte counts.  hps_prof_query
 * synchronises on the recorded events.  Process-wide; not thread-safe with
 * respect to concurrent enable/reset.  No counterpart in the reference. */
int hps_prof_enable(int enable);
int hps_prof_reset(void);
int hps_prof_count(void);
int hps_prof_query(int idx, char* name, int name_len, long* launches, double* ms, double* flops, double* bytes);

#ifdef __cplusplus
}
#endif

#endif /* HPS_MI355X_H */
