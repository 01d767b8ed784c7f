// Internal kernel launchers shared by the translation units of libhps_mi355x.
#pragma once

#include "hps_common.h"

namespace hps {

// D[b] = alpha * op(A[b]) @ B[b] + beta * C[b]   (row-major, per batch element)
// op(A) = A (M x K, lda) or, with transA, A^H where A is stored K x M (lda);
// B: K x N (ldb), C/D: M x N (ldc/ldd).  s* are batch strides (0 broadcasts).
// C may be null when beta == 0.  C and D may alias.
template <typename T>
struct GemmArgs {
  int M = 0, N = 0, K = 0;
  double alpha = 1.0, beta = 0.0;
  const T* A = nullptr;
  long lda = 0, sA = 0;
  const T* B = nullptr;
  long ldb = 0, sB = 0;
  const T* C = nullptr;
  long ldc = 0, sC = 0;
  T* D = nullptr;
  long ldd = 0, sD = 0;
  int batch = 1;
  int transA = 0;
  // triA: A is a packed 32 x 32 diagonal-block inverse used as its unit-lower
  // part (1) or its upper part (2); needs K <= 32.  D may then alias B.
  int triA = 0;
};

template <typename T>
hipError_t gemm(const GemmArgs<T>& a, hipStream_t s);

// The products of row blocks of one A with row blocks of one B: blk holds
// n_blk host-side rectangles (r0, nr, k0, nk); block u is
// D[r0 : r0 + nr] = alpha A[r0 : r0 + nr, k0 : k0 + nk] B[k0 : k0 + nk] + beta C[r0 : r0 + nr]
// with a's N, leading dimensions, batch strides, alpha and beta (a.M and
// a.K are not read).  The blocks must write disjoint rows of D.
template <typename T>
hipError_t gemm_grouped(const GemmArgs<T>& a, int n_blk, const int* blk, hipStream_t s);

// In-place batched LU with partial pivoting of the n x n block of each A[b]
// (row stride lda, batch stride sA): P A = L U.  perm[b*n + i] is the row of
// A that became row i of P A.  dinv[b] (dinv_elems(n) values) receives the
// packed inverses of the 32 x 32 diagonal blocks.  scratch holds
// getrf_scratch_bytes(n, batch) bytes.  info[b] = 0, or j+1 for the first
// column j whose pivot candidates are all exactly zero.
size_t getrf_scratch_bytes(int n, int batch);
size_t dinv_elems(int n);
// nopiv: the tall panels (and every panel of at most 16 matrices) without row
// exchanges; for matrices whose natural-order elimination is stable -- the
// caller checks the factors.  nopiv_all: every register-sized panel without
// row exchanges, checked in the kernel: a matrix with a multiplier above 64
// (or not finite) gets info = -1 and must be refactored with pivoting.
template <typename T>
hipError_t getrf_ws(int n, T* A, long lda, long sA, int* perm, T* dinv, int* scratch, int* info, int batch,
                    hipStream_t s, bool nopiv = false, bool nopiv_all = false);

// X = A^{-1} R from getrf_ws's factors.  R (n x nrhs, row stride ldr, batch
// stride sR) is read only; X must not alias R.
template <typename T>
hipError_t getrs_perm(int n, int nrhs, const T* LU, long lda, long sA, const int* perm, const T* dinv, const T* R,
                      long ldr, long sR, T* X, long ldx, long sX, int batch, hipStream_t s);

// X = A^{-1} P^T X in place, X already gathered by the permutation
// (X[i] = R[perm[i]]): the triangular solves of getrs_perm's blocked path.
template <typename T>
hipError_t getrs_inplace(int n, int nrhs, const T* LU, long lda, long sA, const T* dinv, T* X, long ldx, long sX,
                         int batch, hipStream_t s);

// Many-column solves of large low-precision factors through the inverses of
// their 512 x 512 diagonal blocks (products with K = 512 instead of 32-row
// steps).  Eligible: n >= LU_WIDE_MIN, n a multiple of 512, more than 16
// columns; wsp holds blockinv_workspace_elems(n, nrhs) elements (otherwise,
// or with wsp == nullptr, this is getrs_perm).
constexpr int LU_WIDE_MIN = 3072;
constexpr int BLOCKINV_NB = 512;
constexpr int BLOCKINV_MAX_BATCH = 2;  // the inverses are formed matrix by matrix: top levels only
inline bool blockinv_eligible(int n, int nrhs) { return n >= LU_WIDE_MIN && n % BLOCKINV_NB == 0 && nrhs > 16; }
inline size_t blockinv_workspace_elems(int n, int nrhs) {
  return blockinv_eligible(n, nrhs) ? 2 * (size_t)n * BLOCKINV_NB + (size_t)BLOCKINV_NB * nrhs : 0;
}
template <typename T>
hipError_t getrs_perm_blockinv(int n, int nrhs, const T* LU, long lda, long sA, const int* perm, const T* dinv,
                               const T* R, long ldr, long sR, T* X, long ldx, long sX, int batch, T* wsp,
                               hipStream_t s);

// Few-column solves (at most 16 columns) of matrices too large for one
// workgroup, in fewer than 128 matrices: the refinement's up passes through
// the top merge levels' low-precision factors.  form_fewinv writes the
// inverses of the FEWINV_NB x FEWINV_NB diagonal blocks of L (unit lower) and
// U of every matrix: inv holds batch * fewinv_elems(n) elements (all the L
// block inverses, then all the U ones), ws fewinv_workspace_elems(n, batch)
// elements.  getrs_perm_fewinv solves X = A^{-1} R with them (Y: n x nrhs
// scratch per matrix, row stride nrhs); R, X and Y must not alias.  Eligible:
// n > 256 and a multiple of 32.
constexpr int FEWINV_NB = 256;
inline bool fewinv_eligible(int n) { return n > 256 && n % 32 == 0; }
inline int fewinv_blocks(int n) { return (n + FEWINV_NB - 1) / FEWINV_NB; }
inline size_t fewinv_elems(int n) {
  return fewinv_eligible(n) ? 2 * (size_t)fewinv_blocks(n) * FEWINV_NB * FEWINV_NB : 0;
}
inline size_t fewinv_workspace_elems(int n, int batch) {
  return fewinv_eligible(n) ? (size_t)batch * fewinv_blocks(n) * (FEWINV_NB * FEWINV_NB + FEWINV_NB * 32) : 0;
}
template <typename T>
hipError_t form_fewinv(int n, const T* LU, long lda, long sA, const T* dinv, T* inv, T* ws, int batch, hipStream_t s);
template <typename T>
hipError_t getrs_perm_fewinv(int n, int nrhs, const T* LU, long lda, long sA, const int* perm, const T* inv,
                             const T* R, long ldr, long sR, T* X, long ldx, long sX, T* Y, int batch, hipStream_t s);

// X = A^{-H} R from getrf_ws's factors (A^H = U^H L^H P).  W is n x nrhs
// scratch per batch element (row stride nrhs); X, W and R must not alias.
template <typename T>
hipError_t getrs_conjt(int n, int nrhs, const T* LU, long lda, long sA, const int* perm, const T* dinv, const T* R,
                       long ldr, long sR, T* X, long ldx, long sX, T* W, int batch, hipStream_t s);

// dst[b][r][c] = alpha * src[b][r][c]; src == nullptr writes zeros; a batch
// stride of 0 on src broadcasts.
template <typename T>
hipError_t add2d(int rows, int cols, double alpha, const T* src, long lds, long ss, T* dst, long ldd, long sd,
                 int batch, hipStream_t s);

template <typename T>
hipError_t copy2d(int rows, int cols, double alpha, const T* src, long lds, long ss, T* dst, long ldd, long sd,
                  int batch, hipStream_t s);

}  // namespace hps
