// Batched dense LU with partial pivoting and the matching solve, gfx950.
//
// getrf: blocked right-looking, panel width NB = 32.
//   1. panel kernel (one workgroup per matrix): factors the (n-k) x nb panel
//      column by column.  The pivot search is a wave-level argmax (shuffle
//      reduction inside each 64-wide wave, then across waves through LDS).
//      The panel is staged in LDS when it fits (leaf-sized matrices), and is
//      worked on in place in global memory (L2-resident) otherwise.
//   2. swap+TRSM kernel: applies the panel's row swaps to every column outside
//      the panel and forms U12 = L11^{-1} A12 (L11 staged in LDS).
//   3. trailing update A22 -= L21 U12 through the batched MFMA GEMM.
//   The kernel also maintains perm[] (row i of P*A is row perm[i] of A), so
//   the solve applies the whole permutation as one gather.
// getrs: X = A^{-1} R out of place: gather rows by perm, blocked unit-lower
//   forward substitution and upper back substitution (LDS-staged diagonal
//   blocks, MFMA GEMM updates between blocks).
#include "hps_kernels.h"

namespace hps {

constexpr int LU_NB = 32;
constexpr size_t LU_PANEL_LDS_MAX = 150 * 1024;

__global__ void fill_int_kernel(int* p, long n, int v) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void iota_kernel(int* perm, int n, long total) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < total) perm[i] = (int)(i % n);
}

template <typename R>
__device__ __forceinline__ void argmax_merge(R& bv, int& bi, R v, int i) {
  if (v > bv || (v == bv && i < bi)) {
    bv = v;
    bi = i;
  }
}

// One workgroup per batch element.  P points at A[k][k]; panel has m rows and
// nb columns with row stride ldp (LDS: NB, global: lda).
template <typename T, bool USE_LDS, int THREADS>
__global__ __launch_bounds__(THREADS) void getrf_panel_kernel(int n, T* A, long lda, long sA, int k, int nb,
                                                              int* perm, int* ipiv, int* info) {
  using R = typename real_of<T>::type;
  extern __shared__ char smem_raw[];
  __shared__ R red_v[THREADS / 64];
  __shared__ int red_i[THREADS / 64];
  __shared__ int s_piv;
  const int b = blockIdx.x;
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int m = n - k;
  T* Ag = A + (long)b * sA + (long)k * lda + k;
  T* P;
  long ldp;
  if constexpr (USE_LDS) {
    P = reinterpret_cast<T*>(smem_raw);
    ldp = LU_NB;
    for (int idx = tid; idx < m * nb; idx += THREADS) {
      int r = idx / nb, c = idx - r * nb;
      P[r * LU_NB + c] = Ag[(long)r * lda + c];
    }
    __syncthreads();
  } else {
    P = Ag;
    ldp = lda;
  }
  int* permb = perm + (long)b * n;
  int* ipivb = ipiv + (long)b * n;

  for (int j = 0; j < nb; ++j) {
    R bv = R(-1);
    int bi = 0x7fffffff;
    for (int r = j + tid; r < m; r += THREADS) argmax_merge(bv, bi, s_abs1(P[(long)r * ldp + j]), r);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      R ov = __shfl_xor(bv, off, 64);
      int oi = __shfl_xor(bi, off, 64);
      argmax_merge(bv, bi, ov, oi);
    }
    if (lane == 0) {
      red_v[wave] = bv;
      red_i[wave] = bi;
    }
    __syncthreads();
    if (tid == 0) {
      R fv = red_v[0];
      int fi = red_i[0];
      for (int w = 1; w < THREADS / 64; ++w) argmax_merge(fv, fi, red_v[w], red_i[w]);
      // NaN entries never win a comparison; keep the diagonal then so the
      // pivot row always lies inside the panel.
      if (fi < j || fi >= m) fi = j;
      s_piv = fi;
      ipivb[k + j] = k + fi;
      if (fi != j) {
        int t = permb[k + j];
        permb[k + j] = permb[k + fi];
        permb[k + fi] = t;
      }
      if (fv == R(0) && info[b] == 0) info[b] = k + j + 1;
    }
    __syncthreads();
    const int pr = s_piv;
    if (pr != j) {
      for (int c = tid; c < nb; c += THREADS) {
        T t = P[(long)j * ldp + c];
        P[(long)j * ldp + c] = P[(long)pr * ldp + c];
        P[(long)pr * ldp + c] = t;
      }
    }
    __syncthreads();
    T piv = P[(long)j * ldp + j];
    if (!s_is_zero(piv)) {
      T inv = s_inv(piv);
      for (int r = j + 1 + tid; r < m; r += THREADS) {
        T l = s_mul(P[(long)r * ldp + j], inv);
        P[(long)r * ldp + j] = l;
        for (int c = j + 1; c < nb; ++c) P[(long)r * ldp + c] = s_fnms(P[(long)r * ldp + c], l, P[(long)j * ldp + c]);
      }
    }
    __syncthreads();
  }
  if constexpr (USE_LDS) {
    for (int idx = tid; idx < m * nb; idx += THREADS) {
      int r = idx / nb, c = idx - r * nb;
      Ag[(long)r * lda + c] = P[r * LU_NB + c];
    }
  }
}

// Row swaps of panel k applied to all columns outside the panel, and
// U12 = L11^{-1} A12 for the columns right of it.  One thread per column.
template <typename T>
__global__ __launch_bounds__(256) void getrf_swap_trsm_kernel(int n, T* A, long lda, long sA, int k, int nb,
                                                              const int* ipiv) {
  __shared__ T L11[LU_NB][LU_NB + 1];
  __shared__ int spiv[LU_NB];
  const int b = blockIdx.y;
  T* Ab = A + (long)b * sA;
  for (int idx = threadIdx.x; idx < nb * nb; idx += blockDim.x) {
    int r = idx / nb, c = idx - r * nb;
    L11[r][c] = Ab[(long)(k + r) * lda + k + c];
  }
  for (int j = threadIdx.x; j < nb; j += blockDim.x) spiv[j] = ipiv[(long)b * n + k + j];
  __syncthreads();
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  int c = t < k ? t : t + nb;
  if (c >= n) return;
  for (int j = 0; j < nb; ++j) {
    int pr = spiv[j];
    if (pr != k + j) {
      T x = Ab[(long)(k + j) * lda + c];
      Ab[(long)(k + j) * lda + c] = Ab[(long)pr * lda + c];
      Ab[(long)pr * lda + c] = x;
    }
  }
  if (c >= k + nb) {
    T x[LU_NB];
#pragma unroll
    for (int j = 0; j < LU_NB; ++j)
      if (j < nb) x[j] = Ab[(long)(k + j) * lda + c];
#pragma unroll
    for (int j = 0; j < LU_NB; ++j) {
      if (j < nb) {
        for (int i = 0; i < j; ++i) x[j] = s_fnms(x[j], L11[j][i], x[i]);
        Ab[(long)(k + j) * lda + c] = x[j];
      }
    }
  }
}

template <typename T>
static hipError_t launch_panel(int n, T* A, long lda, long sA, int k, int nb, int* perm, int* ipiv, int* info,
                               int batch, hipStream_t s) {
  int m = n - k;
  size_t lds = (size_t)m * LU_NB * sizeof(T);
  if (lds <= LU_PANEL_LDS_MAX) {
    static bool attr_set = false;
    if (!attr_set) {
      hipFuncSetAttribute((const void*)getrf_panel_kernel<T, true, 256>,
                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)LU_PANEL_LDS_MAX);
      attr_set = true;
    }
    hipLaunchKernelGGL((getrf_panel_kernel<T, true, 256>), dim3(batch), dim3(256), lds, s, n, A, lda, sA, k, nb,
                       perm, ipiv, info);
  } else {
    hipLaunchKernelGGL((getrf_panel_kernel<T, false, 1024>), dim3(batch), dim3(1024), 0, s, n, A, lda, sA, k, nb,
                       perm, ipiv, info);
  }
  return hipGetLastError();
}

// perm/info outputs; ipiv is scratch of batch*n ints.
template <typename T>
hipError_t getrf_ws(int n, T* A, long lda, long sA, int* perm, int* ipiv, int* info, int batch, hipStream_t s) {
  if (n <= 0 || batch <= 0) return hipSuccess;
  long total = (long)batch * n;
  hipLaunchKernelGGL(iota_kernel, dim3(ceil_div(total, 256)), dim3(256), 0, s, perm, n, total);
  hipLaunchKernelGGL(fill_int_kernel, dim3(ceil_div(batch, 256)), dim3(256), 0, s, info, (long)batch, 0);
  for (int k = 0; k < n; k += LU_NB) {
    int nb = n - k < LU_NB ? n - k : LU_NB;
    hipError_t e = launch_panel<T>(n, A, lda, sA, k, nb, perm, ipiv, info, batch, s);
    if (e != hipSuccess) return e;
    int others = n - nb;
    if (others > 0) {
      dim3 grid(ceil_div(others, 256), batch);
      hipLaunchKernelGGL((getrf_swap_trsm_kernel<T>), grid, dim3(256), 0, s, n, A, lda, sA, k, nb, ipiv);
    }
    int rem = n - k - nb;
    if (rem > 0) {
      GemmArgs<T> g;
      g.M = rem;
      g.N = rem;
      g.K = nb;
      g.alpha = -1.0;
      g.beta = 1.0;
      g.A = A + (long)(k + nb) * lda + k;
      g.lda = lda;
      g.sA = sA;
      g.B = A + (long)k * lda + k + nb;
      g.ldb = lda;
      g.sB = sA;
      g.C = A + (long)(k + nb) * lda + k + nb;
      g.ldc = lda;
      g.sC = sA;
      g.D = const_cast<T*>(g.C);
      g.ldd = lda;
      g.sD = sA;
      g.batch = batch;
      e = gemm<T>(g, s);
      if (e != hipSuccess) return e;
    }
  }
  return hipGetLastError();
}

// X[b][i][c] = R[b][perm[b][i]][c]
template <typename T>
__global__ void gather_rows_kernel(int n, int nrhs, const int* perm, const T* R, long ldr, long sR, T* X, long ldx,
                                   long sX) {
  const int b = blockIdx.z;
  const int i = blockIdx.y;
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nrhs) return;
  int src = perm[(long)b * n + i];
  X[(long)b * sX + (long)i * ldx + c] = R[(long)b * sR + (long)src * ldr + c];
}

template <typename T, bool UPPER>
__global__ __launch_bounds__(256) void trsm_block_kernel(int nrhs, const T* LU, long lda, long sA, T* X, long ldx,
                                                         long sX, int k, int nb) {
  __shared__ T Ls[LU_NB][LU_NB + 1];
  const int b = blockIdx.y;
  const T* Lb = LU + (long)b * sA;
  for (int idx = threadIdx.x; idx < nb * nb; idx += blockDim.x) {
    int r = idx / nb, c = idx - r * nb;
    Ls[r][c] = Lb[(long)(k + r) * lda + k + c];
  }
  __syncthreads();
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nrhs) return;
  T* Xb = X + (long)b * sX;
  T x[LU_NB];
#pragma unroll
  for (int j = 0; j < LU_NB; ++j)
    if (j < nb) x[j] = Xb[(long)(k + j) * ldx + c];
  if constexpr (!UPPER) {
#pragma unroll
    for (int j = 0; j < LU_NB; ++j)
      if (j < nb)
        for (int i = 0; i < j; ++i) x[j] = s_fnms(x[j], Ls[j][i], x[i]);
  } else {
#pragma unroll
    for (int jj = LU_NB - 1; jj >= 0; --jj) {
      if (jj < nb) {
        for (int i = jj + 1; i < nb; ++i) x[jj] = s_fnms(x[jj], Ls[jj][i], x[i]);
        x[jj] = s_mul(x[jj], s_inv(Ls[jj][jj]));
      }
    }
  }
#pragma unroll
  for (int j = 0; j < LU_NB; ++j)
    if (j < nb) Xb[(long)(k + j) * ldx + c] = x[j];
}

template <typename T>
hipError_t getrs_perm(int n, int nrhs, const T* LU, long lda, long sA, const int* perm, const T* R, long ldr,
                      long sR, T* X, long ldx, long sX, int batch, hipStream_t s) {
  if (n <= 0 || nrhs <= 0 || batch <= 0) return hipSuccess;
  // batch goes in grid.z (<= 65535 in practice for this library's shapes)
  for (int b0 = 0; b0 < batch; b0 += 65535) {
    int bc = batch - b0 < 65535 ? batch - b0 : 65535;
    dim3 grid(ceil_div(nrhs, 64), n, bc);
    hipLaunchKernelGGL((gather_rows_kernel<T>), grid, dim3(64), 0, s, n, nrhs, perm + (long)b0 * n,
                       R + (long)b0 * sR, ldr, sR, X + (long)b0 * sX, ldx, sX);
  }
  dim3 tgrid(ceil_div(nrhs, 256), batch);
  for (int k = 0; k < n; k += LU_NB) {
    int nb = n - k < LU_NB ? n - k : LU_NB;
    hipLaunchKernelGGL((trsm_block_kernel<T, false>), tgrid, dim3(256), 0, s, nrhs, LU, lda, sA, X, ldx, sX, k, nb);
    int rem = n - k - nb;
    if (rem > 0) {
      GemmArgs<T> g;
      g.M = rem;
      g.N = nrhs;
      g.K = nb;
      g.alpha = -1.0;
      g.beta = 1.0;
      g.A = LU + (long)(k + nb) * lda + k;
      g.lda = lda;
      g.sA = sA;
      g.B = X + (long)k * ldx;
      g.ldb = ldx;
      g.sB = sX;
      g.C = X + (long)(k + nb) * ldx;
      g.ldc = ldx;
      g.sC = sX;
      g.D = X + (long)(k + nb) * ldx;
      g.ldd = ldx;
      g.sD = sX;
      g.batch = batch;
      hipError_t e = gemm<T>(g, s);
      if (e != hipSuccess) return e;
    }
  }
  int nblk = (n + LU_NB - 1) / LU_NB;
  for (int bi = nblk - 1; bi >= 0; --bi) {
    int k = bi * LU_NB;
    int nb = n - k < LU_NB ? n - k : LU_NB;
    hipLaunchKernelGGL((trsm_block_kernel<T, true>), tgrid, dim3(256), 0, s, nrhs, LU, lda, sA, X, ldx, sX, k, nb);
    if (k > 0) {
      GemmArgs<T> g;
      g.M = k;
      g.N = nrhs;
      g.K = nb;
      g.alpha = -1.0;
      g.beta = 1.0;
      g.A = LU + k;
      g.lda = lda;
      g.sA = sA;
      g.B = X + (long)k * ldx;
      g.ldb = ldx;
      g.sB = sX;
      g.C = X;
      g.ldc = ldx;
      g.sC = sX;
      g.D = X;
      g.ldd = ldx;
      g.sD = sX;
      g.batch = batch;
      hipError_t e = gemm<T>(g, s);
      if (e != hipSuccess) return e;
    }
  }
  return hipGetLastError();
}

#define HPS_INST_LU(T)                                                                                        \
  template hipError_t getrf_ws<T>(int, T*, long, long, int*, int*, int*, int, hipStream_t);                   \
  template hipError_t getrs_perm<T>(int, int, const T*, long, long, const int*, const T*, long, long, T*, long, \
                                    long, int, hipStream_t);
HPS_INST_LU(double)
HPS_INST_LU(float)
HPS_INST_LU(c128)
HPS_INST_LU(c64)

}  // namespace hps
