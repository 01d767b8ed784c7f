// C ABI of libhps_mi355x and the HPS stage kernels behind it.
//
// Stage entry points (leaf solve, one merge level, one down-pass level, leaf
// apply) compose the batched LU/GEMM primitives with three kinds of
// memory-bound kernels written here:
//   * leaf operator assembly  A = sum_k diag(c_k) D_k  for a row band,
//   * index-table gathers that build the merge blocks D, C, B, A and the
//     interface/exterior data vectors from the children's T and h,
//   * the down-pass scatter of parent/interface boundary data to children.
// The index tables (which child and local boundary index each merged node
// comes from) are computed once on the host per tree level; they encode the
// reference's block index sets, rolls, flips and row permutation, so one
// merge code path serves 2D DtN, 2D ItI and 3D DtN.
#include <cstdio>
#include <cstring>
#include <string>

#include <mutex>

#include "hps_capi_internal.h"

namespace hps {

// ---------------------------------------------------------------------------
// generic batched 2D copy
// ---------------------------------------------------------------------------
template <typename T>
__global__ void copy2d_kernel(int rows, int cols, double alpha, const T* src, long lds, long ss, T* dst, long ldd,
                              long sd, int batch) {
  // 32-bit index arithmetic (the per-matrix element count stays below 2^31)
  const int total = rows * cols;
  for (int idx = blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += gridDim.x * blockDim.x) {
    const int r = idx / cols, c = idx - r * cols;
    for (int b = blockIdx.y; b < batch; b += gridDim.y) {
      T v = src ? s_scale(src[(long)b * ss + (long)r * lds + c], alpha) : s_zero<T>();
      dst[(long)b * sd + (long)r * ldd + c] = v;
    }
  }
}

template <typename T>
hipError_t copy2d(int rows, int cols, double alpha, const T* src, long lds, long ss, T* dst, long ldd, long sd,
                  int batch, hipStream_t s) {
  if (rows <= 0 || cols <= 0 || batch <= 0) return hipSuccess;
  long total = (long)rows * cols;
  if (total >= (1L << 31)) return hipErrorInvalidValue;  // the kernel indexes one matrix in 32 bits
  unsigned gx = ceil_div(total, 256);
  if (gx > 4096) gx = 4096;
  unsigned gy = batch < 65535 ? batch : 65535;
  HPS_PROF("copy2d", T, 0.0, (src ? 2.0 : 1.0) * sizeof(T) * (double)total * batch);
  hipLaunchKernelGGL((copy2d_kernel<T>), dim3(gx, gy), dim3(256), 0, s, rows, cols, alpha, src, lds, ss, dst, ldd,
                     sd, batch);
  return hipGetLastError();
}

// dst[b][r][c] += alpha * src[b][r][c]
template <typename T>
__global__ void add2d_kernel(int rows, int cols, double alpha, const T* src, long lds, long ss, T* dst, long ldd,
                             long sd, int batch) {
  long total = (long)rows * cols;
  for (long idx = (long)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += (long)gridDim.x * blockDim.x) {
    int r = (int)(idx / cols), c = (int)(idx - (long)r * cols);
    for (int b = blockIdx.y; b < batch; b += gridDim.y) {
      T* d = dst + (long)b * sd + (long)r * ldd + c;
      *d = s_add(*d, s_scale(src[(long)b * ss + (long)r * lds + c], alpha));
    }
  }
}

template <typename T>
hipError_t add2d(int rows, int cols, double alpha, const T* src, long lds, long ss, T* dst, long ldd, long sd,
                 int batch, hipStream_t s) {
  if (rows <= 0 || cols <= 0 || batch <= 0) return hipSuccess;
  long total = (long)rows * cols;
  unsigned gx = ceil_div(total, 256);
  if (gx > 4096) gx = 4096;
  unsigned gy = batch < 65535 ? batch : 65535;
  HPS_PROF("add2d", T, 0.0, 3.0 * sizeof(T) * (double)total * batch);
  hipLaunchKernelGGL((add2d_kernel<T>), dim3(gx, gy), dim3(256), 0, s, rows, cols, alpha, src, lds, ss, dst, ldd, sd,
                     batch);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// leaf operator assembly: out[l][r][c] = sum_i coef_i[l][row0+r] * D_{op_i}[row0+r][col0+c]
// op index < 0 means the identity operator.
// ---------------------------------------------------------------------------
// One workgroup per (operator row, group of ASM_LEAVES consecutive leaves):
// the row's coefficient values are uniform across the workgroup (scalar
// loads, once per workgroup), each thread reads the operator entries of its
// columns (shared by every leaf, so they come from cache) and writes them,
// combined, for each leaf of the group: every store is a coalesced row
// segment.
constexpr int ASM_LEAVES = 4;

template <typename T>
__global__ __launch_bounds__(256) void assemble_kernel(int n, int nl, int ncoef, CoefOps ops_idx, const T* coeffs,
                                                       const T* diffops, int row0, int nrows, int col0, int ncols,
                                                       T* out, long ldo, long so) {
  const int r = blockIdx.x;
  if (r >= nrows) return;
  const int gr = row0 + r;
  for (int l0 = blockIdx.y * ASM_LEAVES; l0 < nl; l0 += gridDim.y * ASM_LEAVES) {
    const int nlg = nl - l0 < ASM_LEAVES ? nl - l0 : ASM_LEAVES;
    for (int c = threadIdx.x; c < ncols; c += blockDim.x) {
      const int gc = col0 + c;
      T acc[ASM_LEAVES];
#pragma unroll
      for (int j = 0; j < ASM_LEAVES; ++j) acc[j] = s_zero<T>();
      for (int i = 0; i < ncoef; ++i) {
        const int op = ops_idx.op[i];
        const T d = op < 0 ? (gr == gc ? s_one<T>() : s_zero<T>()) : diffops[((long)op * n + gr) * n + gc];
        const T* cfp = coeffs + ((long)i * nl + l0) * n + gr;  // uniform across the workgroup
#pragma unroll
        for (int j = 0; j < ASM_LEAVES; ++j)
          if (j < nlg) acc[j] = s_add(acc[j], s_mul(cfp[(long)j * n], d));
      }
#pragma unroll
      for (int j = 0; j < ASM_LEAVES; ++j)
        if (j < nlg) out[(long)(l0 + j) * so + (long)r * ldo + c] = acc[j];
    }
  }
}

// ItI leaf right-hand sides, gathered by the LU's row permutation:
// X[b][i][c] = R[b][perm[b][i]][c] with R = [P 0; 0 f_int] (the scaled
// Gauss-to-Chebyshev map P in the first nb rows and nq columns, the source's
// interior rows in the last nsrc columns), without forming R.
template <typename T>
__global__ __launch_bounds__(256) void iti_rhs_gather_kernel(int n, int nb, int nq, int nsrc, const int* perm,
                                                             const T* P, const T* source, T* X, int nl) {
  const int w = nq + nsrc;
  const int per = n * w;
  for (int b = blockIdx.y; b < nl; b += gridDim.y) {
    const int* pb = perm + (long)b * n;
    T* Xb = X + (long)b * per;
    for (int idx = blockIdx.x * blockDim.x + threadIdx.x; idx < per; idx += gridDim.x * blockDim.x) {
      const int i = idx / w, c = idx - i * w;
      const int j = pb[i];
      T v = s_zero<T>();
      if (c < nq) {
        if (j < nb) v = P[j * nq + c];
      } else if (source && j >= nb) {
        v = source[((long)b * n + j) * nsrc + (c - nq)];
      }
      Xb[idx] = v;
    }
  }
}

// ---------------------------------------------------------------------------
// merge gathers
// rows/cols entries: int4 (c0, l0, c1, l1); c < 0 marks an unused slot.
// out[m][r][c] = sign * sum_{a,b} [rc_a == cc_b] Tin[m*nc + rc_a][rl_a][cl_b] + (add_id && r == c)
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void merge_gather_kernel(const T* Tin, int nc, int nec, const int4* rows, int nr,
                                                           const int4* cols, int ncl, double sign, int add_id,
                                                           T* out, long ldo, long so, int nm) {
  long total = (long)nr * ncl;
  for (long idx = (long)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += (long)gridDim.x * blockDim.x) {
    int r = (int)(idx / ncl), c = (int)(idx - (long)r * ncl);
    int4 re = rows[r];
    int4 ce = cols[c];
    for (int m = blockIdx.y; m < nm; m += gridDim.y) {
      const T* Tm = Tin + (long)m * nc * nec * nec;
      T acc = s_zero<T>();
      if (re.x >= 0) {
        const T* Tc = Tm + (long)re.x * nec * nec + (long)re.y * nec;
        if (ce.x == re.x) acc = s_add(acc, Tc[ce.y]);
        if (ce.z == re.x) acc = s_add(acc, Tc[ce.w]);
      }
      if (re.z >= 0) {
        const T* Tc = Tm + (long)re.z * nec * nec + (long)re.w * nec;
        if (ce.x == re.z) acc = s_add(acc, Tc[ce.y]);
        if (ce.z == re.z) acc = s_add(acc, Tc[ce.w]);
      }
      acc = s_scale(acc, sign);
      if (add_id && r == c) acc = s_add(acc, s_one<T>());
      out[(long)m * so + (long)r * ldo + c] = acc;
    }
  }
}

// out[m][r][s] = sign * sum_a h[m*nc + rc_a][rl_a][s]
template <typename T>
__global__ __launch_bounds__(256) void vec_gather_kernel(const T* h, int nc, int nec, int nsrc, const int4* rows,
                                                         int nr, double sign, T* out, int nm) {
  long total = (long)nr * nsrc;
  for (long idx = (long)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += (long)gridDim.x * blockDim.x) {
    int r = (int)(idx / nsrc), s = (int)(idx - (long)r * nsrc);
    int4 re = rows[r];
    for (int m = blockIdx.y; m < nm; m += gridDim.y) {
      const T* hm = h + (long)m * nc * nec * nsrc;
      T acc = s_zero<T>();
      if (re.x >= 0) acc = s_add(acc, hm[((long)re.x * nec + re.y) * nsrc + s]);
      if (re.z >= 0) acc = s_add(acc, hm[((long)re.z * nec + re.w) * nsrc + s]);
      out[((long)m * nr + r) * nsrc + s] = s_scale(acc, sign);
    }
  }
}

// Down pass scatter: child slot (c, l) takes parent exterior data (src >= 0)
// or interface data (src < 0 -> index -src-1).
template <typename T>
__global__ __launch_bounds__(256) void down_scatter_kernel(int nc, int nec, int ne, int ni, int nsrc,
                                                           const int* child_src, const T* g_par, const T* g_int,
                                                           T* g_out, int nm) {
  long total = (long)nc * nec * nsrc;
  for (long idx = (long)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += (long)gridDim.x * blockDim.x) {
    int slot = (int)(idx / nsrc), s = (int)(idx - (long)slot * nsrc);
    int src = child_src[slot];
    for (int m = blockIdx.y; m < nm; m += gridDim.y) {
      T v = src >= 0 ? (g_par ? g_par[((long)m * ne + src) * nsrc + s] : s_zero<T>())
                     : g_int[((long)m * ni + (-src - 1)) * nsrc + s];
      g_out[((long)m * nc * nec + slot) * nsrc + s] = v;
    }
  }
}

}  // namespace hps

// ===========================================================================
// C ABI
// ===========================================================================
// The entry points below get C linkage from their declarations in
// hps_mi355x.h; the template helpers between them stay C++.

// ---------------------------------------------------------------------------
// Library-owned device scratch for the entry points that take no workspace
// (hps_down_level_static).  One copy per GPU, held until the process ends.
// In host code `scratch` names a host-side stand-in, not device memory: the
// device address of each GPU's copy comes from hipGetSymbolAddress, resolved
// once per device when the first handle on that device is created, and only
// that address is ever handed to kernels.
// ---------------------------------------------------------------------------
#define HPS_MAX_DEVICES 64
static __device__ __align__(256) unsigned char scratch[HPS_STATIC_SCRATCH_BYTES];
static void* g_scratch_ptr[HPS_MAX_DEVICES];
static hipError_t g_scratch_err[HPS_MAX_DEVICES];
static bool g_scratch_done[HPS_MAX_DEVICES];
static std::mutex g_scratch_mu;

// Resolve this device's copy of `scratch` with the device current.
static void resolve_scratch(int dev) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (dev < 0 || dev >= HPS_MAX_DEVICES || g_scratch_done[dev]) return;
  int cur = -1;
  hipError_t e = hipGetDevice(&cur);
  if (e == hipSuccess && cur != dev) e = hipSetDevice(dev);
  void* p = nullptr;
  if (e == hipSuccess) e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (cur >= 0 && cur != dev) (void)hipSetDevice(cur);
  g_scratch_ptr[dev] = e == hipSuccess ? p : nullptr;
  g_scratch_err[dev] = e;
  g_scratch_done[dev] = true;
}

int hps_create(void* stream, hps_handle_t* out) {
  if (!out) return HPS_ERR_ARG;
  hps_handle_s* h = new hps_handle_s();
  h->stream = (hipStream_t)stream;
  int dev = -1;
  hipError_t e = stream ? hipStreamGetDevice((hipStream_t)stream, &dev) : hipGetDevice(&dev);
  h->device = e == hipSuccess ? dev : -1;
  if (h->device >= 0) resolve_scratch(h->device);
  *out = h;
  return HPS_OK;
}

size_t hps_static_scratch_bytes(void) { return HPS_STATIC_SCRATCH_BYTES; }

int hps_set_merge_pivoting(hps_handle_t h, int partial_pivoting) {
  if (!h) return HPS_ERR_ARG;
  h->merge_nopiv = partial_pivoting == 0;
  return HPS_OK;
}

int hps_set_leaf_pivoting(hps_handle_t h, int partial_pivoting) {
  if (!h) return HPS_ERR_ARG;
  h->leaf_nopiv = partial_pivoting == 0;
  return HPS_OK;
}

int hps_destroy(hps_handle_t h) {
  delete h;
  return HPS_OK;
}

int hps_set_stream(hps_handle_t h, void* stream) {
  if (!h) return HPS_ERR_ARG;
  h->stream = (hipStream_t)stream;
  return HPS_OK;
}

const char* hps_last_error(hps_handle_t h) { return h ? h->err.c_str() : "null handle"; }

int hps_version(void) { return HPS_ABI_VERSION; }

// ---------------------------------------------------------------------------
// primitives
// ---------------------------------------------------------------------------
template <typename T>
static int gemm_t(hps_handle_t h, int M, int N, int K, double alpha, const void* A, long lda, long sA, const void* B,
                  long ldb, long sB, double beta, const void* C, long ldc, long sC, void* D, long ldd, long sD,
                  int batch) {
  GemmArgs<T> g;
  g.M = M;
  g.N = N;
  g.K = K;
  g.alpha = alpha;
  g.beta = beta;
  g.A = (const T*)A;
  g.lda = lda;
  g.sA = sA;
  g.B = (const T*)B;
  g.ldb = ldb;
  g.sB = sB;
  g.C = (const T*)C;
  g.ldc = ldc;
  g.sC = sC;
  g.D = (T*)D;
  g.ldd = ldd;
  g.sD = sD;
  g.batch = batch;
  HPS_CHECK_HIP(h, gemm<T>(g, h->stream));
  return HPS_OK;
}

int hps_gemm_batched(hps_handle_t h, int dtype, int M, int N, int K, double alpha, const void* A, long lda, long sA,
                     const void* B, long ldb, long sB, double beta, const void* C, long ldc, long sC, void* D,
                     long ldd, long sD, int batch) {
  if (!h) return HPS_ERR_ARG;
  HPS_ARG(h, M >= 0 && N >= 0 && K >= 0 && batch >= 0, "negative size");
  HPS_ARG(h, D != nullptr && (K == 0 || (A && B)), "null matrix");
  HPS_ARG(h, beta == 0.0 || C != nullptr, "beta != 0 needs C");
  DISPATCH(dtype, gemm_t, h, M, N, K, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, D, ldd, sD, batch);
}

size_t hps_getrf_workspace(int n, int batch) { return align_up(getrf_scratch_bytes(n, batch)); }

size_t hps_dinv_size(int n) { return dinv_elems(n); }

template <typename T>
static int getrf_t(hps_handle_t h, int n, void* A, long lda, long sA, int* perm, void* dinv, int* info, int batch,
                   void* ws) {
  HPS_CHECK_HIP(h, getrf_ws<T>(n, (T*)A, lda, sA, perm, (T*)dinv, (int*)ws, info, batch, h->stream));
  return HPS_OK;
}

int hps_getrf_batched(hps_handle_t h, int dtype, int n, void* A, long lda, long sA, int* perm, void* dinv, int* info,
                      int batch, void* ws, size_t ws_bytes) {
  if (!h) return HPS_ERR_ARG;
  HPS_ARG(h, n >= 0 && n <= 12288 && batch >= 0 && lda >= n, "bad size (0 <= n <= 12288, lda >= n)");
  HPS_ARG(h, A && perm && dinv && info, "null pointer");
  if (ws_bytes < hps_getrf_workspace(n, batch)) {
    h->err = "getrf workspace too small";
    return HPS_ERR_WORKSPACE;
  }
  DISPATCH(dtype, getrf_t, h, n, A, lda, sA, perm, dinv, info, batch, ws);
}

template <typename T>
static int getrs_t(hps_handle_t h, int n, int nrhs, const void* LU, long lda, long sA, const int* perm,
                   const void* dinv, const void* R, long ldr, long sR, void* X, long ldx, long sX, int batch) {
  HPS_CHECK_HIP(h, getrs_perm<T>(n, nrhs, (const T*)LU, lda, sA, perm, (const T*)dinv, (const T*)R, ldr, sR, (T*)X,
                                 ldx, sX, batch, h->stream));
  return HPS_OK;
}

int hps_getrs_batched(hps_handle_t h, int dtype, int n, int nrhs, const void* LU, long lda, long sA, const int* perm,
                      const void* dinv, const void* R, long ldr, long sR, void* X, long ldx, long sX, int batch) {
  if (!h) return HPS_ERR_ARG;
  HPS_ARG(h, n >= 0 && nrhs >= 0 && batch >= 0, "bad size");
  HPS_ARG(h, LU && perm && dinv && R && X, "null pointer");
  HPS_ARG(h, R != X, "getrs is out of place: R and X must differ");
  DISPATCH(dtype, getrs_t, h, n, nrhs, LU, lda, sA, perm, dinv, R, ldr, sR, X, ldx, sX, batch);
}

// ---------------------------------------------------------------------------
// leaf solve
// ---------------------------------------------------------------------------
size_t hps_leaf_workspace(int dtype, int kind, int nl, int n, int nb, int nq, int nsrc) {
  size_t es = dtype_size(dtype);
  int nrows = kind == HPS_KIND_DTN ? n - nb : n;
  size_t rhs = align_up((size_t)nl * nrows * (nq + nsrc) * es);
  return 2 * rhs + align_up(getrf_scratch_bytes(nrows, nl));
}

template <typename T>
static int leaf_t(hps_handle_t h, int kind, int nl, int n, int nb, int nq, int nsrc, int ncoef, const int* coef_ops,
                  const void* coeffs, const void* diffops, const void* Pv, const void* Qv, const void* Gv,
                  const void* source, void* Yv, void* Tv, void* vv, void* hv, void* luv, int* perm, void* dinvv,
                  int* info, void* ws) {
  hipStream_t s = h->stream;
  const T* P = (const T*)Pv;
  const T* Q = (const T*)Qv;
  T* Y = (T*)Yv;
  T* Tm = (T*)Tv;
  T* lu = (T*)luv;
  const int ni = n - nb;
  const int nrows = kind == HPS_KIND_DTN ? ni : n;
  const int w = nq + nsrc;
  size_t es = sizeof(T);
  char* wp = (char*)ws;
  T* R = (T*)wp;
  wp += align_up((size_t)nl * nrows * w * es);
  T* X = (T*)wp;
  wp += align_up((size_t)nl * nrows * w * es);
  int* scratch = (int*)wp;
  T* dinv = (T*)dinvv;

  CoefOps co;
  for (int i = 0; i < 16; ++i) co.op[i] = i < ncoef ? coef_ops[i] : 0;

  // 1. assemble the operator rows the factorisation needs
  const long ld_lu = n;
  const long s_lu = (long)nrows * n;
  int arow0 = nb, an = ni;
  T* aout = kind == HPS_KIND_DTN ? lu : lu + (long)nb * n;
  {
    HPS_PROF("leaf_assemble", T, flop_mult<T>() * 2.0 * ncoef * (double)an * n * nl,
             sizeof(T) * ((double)an * n * nl + (double)ncoef * n * nl));
    dim3 g(an, ceil_div(nl, ASM_LEAVES) < 65535 ? ceil_div(nl, ASM_LEAVES) : 65535);
    hipLaunchKernelGGL((assemble_kernel<T>), g, dim3(256), 0, s, n, nl, ncoef, co, (const T*)coeffs,
                       (const T*)diffops, arow0, an, 0, n, aout, ld_lu, s_lu);
    HPS_CHECK_HIP(h, hipGetLastError());
  }
  if (kind == HPS_KIND_ITI) {
    // incoming-impedance rows G on top of B
    HPS_CHECK_HIP(h, copy2d<T>(nb, n, 1.0, (const T*)Gv, n, 0, lu, ld_lu, s_lu, nl, s));
  }

  // 2. right-hand sides
  if (kind == HPS_KIND_DTN) {
    // R[:, :, :nq] = A_ie P
    GemmArgs<T> g;
    g.M = ni;
    g.N = nq;
    g.K = nb;
    g.A = lu;
    g.lda = ld_lu;
    g.sA = s_lu;
    g.B = P;
    g.ldb = nq;
    g.sB = 0;
    g.D = R;
    g.ldd = w;
    g.sD = (long)ni * w;
    g.batch = nl;
    HPS_CHECK_HIP(h, gemm<T>(g, s));
    if (nsrc > 0)
      HPS_CHECK_HIP(h, copy2d<T>(ni, nsrc, 1.0, source ? (const T*)source + (long)nb * nsrc : nullptr, nsrc,
                                 (long)n * nsrc, R + nq, w, (long)ni * w, nl, s));
  }

  // 3. factor and solve
  T* Afac = kind == HPS_KIND_DTN ? lu + nb : lu;
  HPS_CHECK_HIP(h, getrf_ws<T>(nrows, Afac, ld_lu, s_lu, perm, dinv, scratch, info, nl, s, false, h->leaf_nopiv));
  if (kind == HPS_KIND_DTN) {
    HPS_CHECK_HIP(h, getrs_perm<T>(nrows, w, Afac, ld_lu, s_lu, perm, dinv, R, w, (long)nrows * w, X, w,
                                   (long)nrows * w, nl, s));
  } else {
    // ItI: the right-hand sides [P 0; 0 f_int] go straight to X in the
    // factors' row order (no R), then the solves run in place
    {
      HPS_PROF("leaf_rhs_gather", T, 0.0, sizeof(T) * (double)n * w * nl);
      const int per = n * w;
      dim3 g(ceil_div(per, 256) < 64 ? ceil_div(per, 256) : 64, nl < 65535 ? nl : 65535);
      hipLaunchKernelGGL((iti_rhs_gather_kernel<T>), g, dim3(256), 0, s, n, nb, nq, nsrc, perm, P,
                         (const T*)source, X, nl);
      HPS_CHECK_HIP(h, hipGetLastError());
    }
    HPS_CHECK_HIP(h, getrs_inplace<T>(nrows, w, Afac, ld_lu, s_lu, dinv, X, w, (long)nrows * w, nl, s));
  }

  // 4. unpack Y and v
  if (kind == HPS_KIND_DTN) {
    HPS_CHECK_HIP(h, copy2d<T>(nb, nq, 1.0, P, nq, 0, Y, nq, (long)n * nq, nl, s));
    HPS_CHECK_HIP(h, copy2d<T>(ni, nq, -1.0, X, w, (long)ni * w, Y + (long)nb * nq, nq, (long)n * nq, nl, s));
    if (nsrc > 0) {
      HPS_CHECK_HIP(h, copy2d<T>(nb, nsrc, 1.0, nullptr, 0, 0, (T*)vv, nsrc, (long)n * nsrc, nl, s));
      HPS_CHECK_HIP(h, copy2d<T>(ni, nsrc, 1.0, X + nq, w, (long)ni * w, (T*)vv + (long)nb * nsrc, nsrc,
                                 (long)n * nsrc, nl, s));
    }
  } else {
    HPS_CHECK_HIP(h, copy2d<T>(n, nq, 1.0, X, w, (long)n * w, Y, nq, (long)n * nq, nl, s));
    if (nsrc > 0)
      HPS_CHECK_HIP(h, copy2d<T>(n, nsrc, 1.0, X + nq, w, (long)n * w, (T*)vv, nsrc, (long)n * nsrc, nl, s));
  }

  // 5. T = Q Y, h = Q v
  {
    GemmArgs<T> g;
    g.M = nq;
    g.N = nq;
    g.K = n;
    g.A = Q;
    g.lda = n;
    g.sA = 0;
    g.B = Y;
    g.ldb = nq;
    g.sB = (long)n * nq;
    g.D = Tm;
    g.ldd = nq;
    g.sD = (long)nq * nq;
    g.batch = nl;
    HPS_CHECK_HIP(h, gemm<T>(g, s));
  }
  if (nsrc > 0) {
    GemmArgs<T> g;
    g.M = nq;
    g.N = nsrc;
    g.K = n;
    g.A = Q;
    g.lda = n;
    g.sA = 0;
    g.B = (const T*)vv;
    g.ldb = nsrc;
    g.sB = (long)n * nsrc;
    g.D = (T*)hv;
    g.ldd = nsrc;
    g.sD = (long)nq * nsrc;
    g.batch = nl;
    HPS_CHECK_HIP(h, gemm<T>(g, s));
  }
  return HPS_OK;
}

int hps_leaf_solve(hps_handle_t h, int dtype, int kind, int nl, int n, int nb, int nq, int nsrc, int ncoef,
                   const int* coef_ops, const void* coeffs, const void* diffops, const void* P, const void* Q,
                   const void* G, const void* source, void* Y, void* T, void* v, void* hvec, void* lu, int* perm,
                   void* dinv, int* info, void* ws, size_t ws_bytes) {
  if (!h) return HPS_ERR_ARG;
  HPS_ARG(h, kind == HPS_KIND_DTN || kind == HPS_KIND_ITI, "kind must be DtN or ItI");
  HPS_ARG(h, nl > 0 && n > nb && nb > 0 && nq > 0 && nsrc >= 0, "bad leaf sizes");
  HPS_ARG(h, ncoef > 0 && ncoef <= 16 && coef_ops, "1..16 coefficients required");
  HPS_ARG(h, coeffs && diffops && P && Q && Y && T && lu && perm && dinv && info, "null pointer");
  HPS_ARG(h, kind == HPS_KIND_DTN || G, "ItI needs G");
  HPS_ARG(h, nsrc == 0 || (v && hvec), "v/h required when nsrc > 0");
  if (ws_bytes < hps_leaf_workspace(dtype, kind, nl, n, nb, nq, nsrc)) {
    h->err = "leaf workspace too small";
    return HPS_ERR_WORKSPACE;
  }
  DISPATCH(dtype, leaf_t, h, kind, nl, n, nb, nq, nsrc, ncoef, coef_ops, coeffs, diffops, P, Q, G, source, Y, T, v,
           hvec, lu, perm, dinv, info, ws);
}

// ---------------------------------------------------------------------------
// merge level
// ---------------------------------------------------------------------------
// Workspace layout (bytes, each region 256-aligned):
//   Y   ni x ne      right-hand sides -C (internal unknown order), solutions
//   Z   nF x ne      ItI: Y_F - D12 Y_G
//   Yh  ni x nsrc    right-hand sides -h_int
//   Zh  nF x nsrc
//   D12, D21 nF x nF ItI blocks (when the caller does not keep them)
//   B   ne x ni      (when the caller does not keep it)
//   getrf scratch
// Low-precision merge levels of fewer than 128 matrices of more than 256
// unknowns get inverses of their factors' diagonal blocks: the merge's own
// few-column source solve (1 .. 16 columns) goes through them, and a caller
// that passes F_inv_out keeps them for its up passes.  The workspace holds
// them (and the forming and solve scratch) whenever the level qualifies.
static bool merge_block_inverses(int dtype, int nfac, int nm) {
  return (dtype == HPS_F32 || dtype == HPS_C64) && nm < 128 && fewinv_eligible(nfac);
}

size_t hps_merge_workspace(int dtype, int kind, int nm, int ni, int ne, int nsrc) {
  size_t es = dtype_size(dtype);
  const int nF = kind == HPS_KIND_ITI ? ni / 2 : ni;
  const int nfac = kind == HPS_KIND_ITI ? nF : ni;
  const int ns = nsrc > 0 ? nsrc : 1;
  size_t tot = 0;
  tot += align_up((size_t)nm * ni * ne * es);
  tot += kind == HPS_KIND_ITI ? align_up((size_t)nm * nF * ne * es) : 0;
  tot += align_up((size_t)nm * ni * ns * es);
  tot += kind == HPS_KIND_ITI ? align_up((size_t)nm * nF * ns * es) : 0;
  tot += kind == HPS_KIND_ITI ? 2 * align_up((size_t)nm * nF * nF * es) : 0;
  tot += align_up((size_t)nm * ne * ni * es);
  tot += align_up(getrf_scratch_bytes(nfac, nm));
  if ((dtype == HPS_F32 || dtype == HPS_C64) && nm <= BLOCKINV_MAX_BATCH)
    tot += align_up(blockinv_workspace_elems(nfac, ne) * es);
  if (merge_block_inverses(dtype, nfac, nm))
    tot += align_up(nm * fewinv_elems(nfac) * es) + align_up(fewinv_workspace_elems(nfac, nm) * es) +
           align_up((size_t)nm * nfac * ns * es);
  return tot;
}

namespace hps {

template <typename T>
__global__ void add_identity_kernel(T* A, int n, long sA, int nm) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int m = blockIdx.y;
  if (i < n && m < nm) A[(long)m * sA + (long)i * n + i] = s_add(A[(long)m * sA + (long)i * n + i], s_one<T>());
}

// dst[m][map[i]][c] = src[m][i][c]  (row scatter from the internal to the final unknown order)
template <typename T>
__global__ void row_scatter_kernel(const T* src, T* dst, const int* map, int rows, int cols, int nm) {
  long total = (long)rows * cols;
  for (long idx = (long)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += (long)gridDim.x * blockDim.x) {
    int r = (int)(idx / cols), c = (int)(idx - (long)r * cols);
    for (int m = blockIdx.y; m < nm; m += gridDim.y)
      dst[((long)m * rows + map[r]) * cols + c] = src[((long)m * rows + r) * cols + c];
  }
}

// dst[m][r][c] = src[m][map[r]][c]
template <typename T>
__global__ void row_gather_kernel(const T* src, T* dst, const int* map, int rows, int cols, int nm) {
  long total = (long)rows * cols;
  for (long idx = (long)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += (long)gridDim.x * blockDim.x) {
    int r = (int)(idx / cols), c = (int)(idx - (long)r * cols);
    for (int m = blockIdx.y; m < nm; m += gridDim.y)
      dst[((long)m * rows + r) * cols + c] = src[((long)m * rows + map[r]) * cols + c];
  }
}

// grad[k][l][i] = -sum_s Re(conj(lam[l][i][s]) * x_k[l][i][s]) on the rows i >= row0 that the
// coefficients enter (x_k = op_k u from R, or u for the identity), 0 on the rows below
template <typename T>
__global__ __launch_bounds__(256) void coef_grad_kernel(int nl, int n, int row0, int w, int ncoef, CoefOps ops,
                                                        const T* R, int nrow_R, const T* u, const T* lam,
                                                        double* grad) {
  long total = (long)nl * n;
  for (long idx = (long)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += (long)gridDim.x * blockDim.x) {
    const int l = (int)(idx / n), i = (int)(idx - (long)l * n);
    for (int k = 0; k < ncoef; ++k) {
      double g = 0.0;
      if (i >= row0) {
        const int op = ops.op[k];
        for (int sidx = 0; sidx < w; ++sidx) {
          const T x = op < 0 ? u[((long)l * n + i) * w + sidx] : R[((long)l * nrow_R + (long)op * n + i) * w + sidx];
          const T lv = lam[((long)l * n + i) * w + sidx];
          g -= (double)s_re(lv) * (double)s_re(x) + (double)s_im(lv) * (double)s_im(x);
        }
      }
      grad[((long)k * nl + l) * n + i] = g;
    }
  }
}

template <typename T>
static hipError_t gather_launch(hipStream_t s, const char* label, const T* Tin, int nc, int nec, const int4* rows,
                                int nr, const int4* cols, int ncl, double sign, int add_id, T* out, long ldo, long so,
                                int nm) {
  static const std::string nm_ = std::string("merge_gather<") + tname<T>() + ">";
  prof::Scope ps(s, nm_.c_str(), 0.0, 2.0 * sizeof(T) * (double)nr * ncl * nm);
  hipLaunchKernelGGL((merge_gather_kernel<T>), grid_for((long)nr * ncl, nm), dim3(256), 0, s, Tin, nc, nec, rows, nr,
                     cols, ncl, sign, add_id, out, ldo, so, nm);
  return hipGetLastError();
}

template <typename T>
static hipError_t gemm_call(hipStream_t s, int M, int N, int K, double alpha, const T* A, long lda, long sA,
                            const T* B, long ldb, long sB, double beta, const T* C, long ldc, long sC, T* D, long ldd,
                            long sD, int batch) {
  GemmArgs<T> g;
  g.M = M;
  g.N = N;
  g.K = K;
  g.alpha = alpha;
  g.beta = beta;
  g.A = A;
  g.lda = lda;
  g.sA = sA;
  g.B = B;
  g.ldb = ldb;
  g.sB = sB;
  g.C = C;
  g.ldc = ldc;
  g.sC = sC;
  g.D = D;
  g.ldd = ldd;
  g.sD = sD;
  g.batch = batch;
  return gemm<T>(g, s);
}

// The row-block products of a block list (r0, nr, k0, nk) in one grouped
// launch: D[r0:] = alpha A[r0:, k0:k0+nk] B[k0:k0+nk] + beta C[r0:].
template <typename T>
static hipError_t block_gemm(hipStream_t s, int n_blk, const int* blk, int N, double alpha, const T* A, long lda,
                             long sA, const T* B, long ldb, long sB, double beta, const T* C, long ldc, long sC, T* D,
                             long ldd, long sD, int batch) {
  GemmArgs<T> g;
  g.N = N;
  g.alpha = alpha;
  g.beta = beta;
  g.A = A;
  g.lda = lda;
  g.sA = sA;
  g.B = B;
  g.ldb = ldb;
  g.sB = sB;
  g.C = C;
  g.ldc = ldc;
  g.sC = sC;
  g.D = D;
  g.ldd = ldd;
  g.sD = sD;
  g.batch = batch;
  return gemm_grouped<T>(g, n_blk, blk, s);
}

}  // namespace hps

// Products with the block-sparse ItI couplings: D = I + [[0, D12], [D21, 0]]
// where row block i of D12 (D21) is nonzero only in one contiguous run of
// column blocks.  blk lists (r0, nr, k0, nk) rectangles that cover every row
// exactly once; n_blk = 0 means dense.  Out[r] = C[r] + alpha Dxy[r, k] X[k].
template <typename T>
static hipError_t coupling_gemm(hipStream_t s, int n_blk, const int* blk, int nF, int w, double alpha, const T* Dxy,
                                long sD, const T* X, long ldx, long sX, const T* C, long ldc, long sC, T* Out,
                                long ldo, long sO, int nm) {
  if (n_blk <= 0)
    return gemm_call<T>(s, nF, w, nF, alpha, Dxy, nF, sD, X, ldx, sX, C ? 1.0 : 0.0, C, ldc, sC, Out, ldo, sO, nm);
  return block_gemm<T>(s, n_blk, blk, w, alpha, Dxy, nF, sD, X, ldx, sX, C ? 1.0 : 0.0, C, ldc, sC, Out, ldo, sO, nm);
}

template <typename T>
static int merge_t(hps_handle_t h, int kind, int nm, int nc, int nec, int ni, int ne, int nsrc, const void* T_in,
                   const void* h_in, const int* int_rows, const int* int_cols, const int* ext,
                   const int* cols_final, const int* final_of_internal, int n_upd, const int* upd, int n_d12,
                   const int* d12_blk, int n_d21, const int* d21_blk, void* S_out, void* T_out, void* h_out,
                   void* gt_out, void* F_lu, int* perm, void* F_dinv, int* info, void* D12_out, void* D21_out,
                   void* B_out, void* F_inv_out, void* ws) {
  hipStream_t s = h->stream;
  const T* Tin = (const T*)T_in;
  const int4* rows = (const int4*)int_rows;
  const int4* cols = (const int4*)int_cols;
  const int4* ex = (const int4*)ext;
  const int4* colsF = (const int4*)(cols_final ? cols_final : int_cols);
  const bool iti = kind == HPS_KIND_ITI;
  const int nF = iti ? ni / 2 : ni;
  const int ns = nsrc > 0 ? nsrc : 1;
  const size_t es = sizeof(T);
  char* wp = (char*)ws;
  auto take = [&](size_t bytes) {
    char* p = wp;
    wp += align_up(bytes);
    return p;
  };
  T* Y = (T*)take((size_t)nm * ni * ne * es);
  T* Z = iti ? (T*)take((size_t)nm * nF * ne * es) : nullptr;
  T* Yh = (T*)take((size_t)nm * ni * ns * es);
  T* Zh = iti ? (T*)take((size_t)nm * nF * ns * es) : nullptr;
  T* D12w = iti ? (T*)take((size_t)nm * nF * nF * es) : nullptr;
  T* D21w = iti ? (T*)take((size_t)nm * nF * nF * es) : nullptr;
  T* Bw = (T*)take((size_t)nm * ne * ni * es);
  constexpr bool lowp = std::is_same<T, float>::value || std::is_same<T, c64>::value;
  const int nfac = iti ? nF : ni;
  T* binv = lowp && nm <= BLOCKINV_MAX_BATCH && blockinv_workspace_elems(nfac, ne)
                ? (T*)take(blockinv_workspace_elems(nfac, ne) * es)
                : nullptr;
  // block inverses (see merge_block_inverses): into F_inv_out when the caller
  // keeps them, else into the workspace when the source solve uses them
  const bool inv_ok = lowp && nm < 128 && fewinv_eligible(nfac);  // merge_block_inverses for this type
  T* finv_ws = inv_ok ? (T*)take(nm * fewinv_elems(nfac) * es) : nullptr;
  T* fws = inv_ok ? (T*)take(fewinv_workspace_elems(nfac, nm) * es) : nullptr;
  T* fsc = inv_ok ? (T*)take((size_t)nm * nfac * ns * es) : nullptr;
  const bool few = inv_ok && nsrc >= 1 && nsrc <= 16;  // the source solve goes through them
  const bool form_inv = inv_ok && (few || F_inv_out);
  T* finv = F_inv_out && inv_ok ? (T*)F_inv_out : finv_ws;
  int* scratch = (int*)wp;
  T* Fl = (T*)F_lu;
  T* dinv = (T*)F_dinv;
  T* D12 = D12_out ? (T*)D12_out : D12w;
  T* D21 = D21_out ? (T*)D21_out : D21w;
  T* Bm = B_out ? (T*)B_out : Bw;
  const long sNF2 = (long)nF * nF;

  // ---- factor ----------------------------------------------------------
  if (!iti) {
    HPS_CHECK_HIP(h, gather_launch<T>(s, "D", Tin, nc, nec, rows, ni, cols, ni, 1.0, 0, Fl, ni, (long)ni * ni, nm));
  } else {
    HPS_CHECK_HIP(h, gather_launch<T>(s, "D12", Tin, nc, nec, rows, nF, cols + nF, nF, 1.0, 0, D12, nF, sNF2, nm));
    HPS_CHECK_HIP(h, gather_launch<T>(s, "D21", Tin, nc, nec, rows + nF, nF, cols, nF, 1.0, 0, D21, nF, sNF2, nm));
    // W = I - D12 D21
    HPS_CHECK_HIP(h, coupling_gemm<T>(s, n_d12, d12_blk, nF, nF, -1.0, D12, sNF2, D21, nF, sNF2, nullptr, 0, 0, Fl, nF,
                                      sNF2, nm));
    hipLaunchKernelGGL((add_identity_kernel<T>), dim3(ceil_div(nF, 256), nm), dim3(256), 0, s, Fl, nF, sNF2, nm);
    HPS_CHECK_HIP(h, hipGetLastError());
  }
  HPS_CHECK_HIP(h, getrf_ws<T>(nF, Fl, nF, sNF2, perm, dinv, scratch, info, nm, s, h->merge_nopiv, h->merge_nopiv));
  if (form_inv) HPS_CHECK_HIP(h, form_fewinv<T>(nF, Fl, nF, sNF2, dinv, finv, fws, nm, s));

  // ---- solves: X = -D^{-1} [C | h_int] -------------------------------------
  // one pass for the C columns (width ne) and one for h (width nsrc)
  for (int pass = 0; pass < 2; ++pass) {
    const int w = pass == 0 ? ne : nsrc;
    if (w <= 0) continue;
    T* Yp = pass == 0 ? Y : Yh;
    T* Zp = pass == 0 ? Z : Zh;
    T* out = (T*)(pass == 0 ? S_out : gt_out);
    if (pass == 0) {
      HPS_CHECK_HIP(h, gather_launch<T>(s, "C", Tin, nc, nec, rows, ni, ex, ne, -1.0, 0, Yp, ne, (long)ni * ne, nm));
    } else {
      static const std::string nmh = std::string("merge_gather_h<") + tname<T>() + ">";
      prof::Scope ps(s, nmh.c_str(), 0.0, 2.0 * es * (double)ni * nsrc * nm);
      hipLaunchKernelGGL((vec_gather_kernel<T>), grid_for((long)ni * nsrc, nm), dim3(256), 0, s, (const T*)h_in, nc,
                         nec, nsrc, rows, ni, -1.0, Yp, nm);
      HPS_CHECK_HIP(h, hipGetLastError());
    }
    const long sY = (long)ni * w;
    if (!iti) {
      if (pass == 1 && few)
        HPS_CHECK_HIP(h, getrs_perm_fewinv<T>(ni, w, Fl, ni, (long)ni * ni, perm, finv, Yp, w, sY, out, w, sY, fsc,
                                              nm, s));
      else
        HPS_CHECK_HIP(h, getrs_perm_blockinv<T>(ni, w, Fl, ni, (long)ni * ni, perm, dinv, Yp, w, sY, out, w, sY, nm,
                                                pass == 0 ? binv : nullptr, s));
    } else {
      const long sZ = (long)nF * w;
      // Z = Y_F - D12 Y_G
      HPS_CHECK_HIP(h, coupling_gemm<T>(s, n_d12, d12_blk, nF, w, -1.0, D12, sNF2, Yp + (long)nF * w, w, sY, Yp, w, sY,
                                        Zp, w, sZ, nm));
      // X_F = W^{-1} Z  (into Y_F)
      if (pass == 1 && few)
        HPS_CHECK_HIP(h, getrs_perm_fewinv<T>(nF, w, Fl, nF, sNF2, perm, finv, Zp, w, sZ, Yp, w, sY, fsc, nm, s));
      else
        HPS_CHECK_HIP(h, getrs_perm_blockinv<T>(nF, w, Fl, nF, sNF2, perm, dinv, Zp, w, sZ, Yp, w, sY, nm,
                                                pass == 0 ? binv : nullptr, s));
      // X_G = Y_G - D21 X_F  (in place)
      HPS_CHECK_HIP(h, coupling_gemm<T>(s, n_d21, d21_blk, nF, w, -1.0, D21, sNF2, Yp, w, sY, Yp + (long)nF * w, w,
                                        sY, Yp + (long)nF * w, w, sY, nm));
      // internal -> final unknown order
      static const std::string nms = std::string("row_scatter<") + tname<T>() + ">";
      prof::Scope ps(s, nms.c_str(), 0.0, 2.0 * es * (double)ni * w * nm);
      hipLaunchKernelGGL((row_scatter_kernel<T>), grid_for((long)ni * w, nm), dim3(256), 0, s, Yp, out,
                         final_of_internal, ni, w, nm);
      HPS_CHECK_HIP(h, hipGetLastError());
    }
  }

  // ---- T = A + B S, h_out = h_ext + B g_tilde (block-sparse B) ---------------
  if (T_out || (nsrc > 0 && h_out) || B_out) {
    HPS_CHECK_HIP(h, gather_launch<T>(s, "B", Tin, nc, nec, ex, ne, colsF, ni, 1.0, 0, Bm, ni, (long)ne * ni, nm));
  }
  for (int pass = 0; pass < 2; ++pass) {
    const int w = pass == 0 ? ne : nsrc;
    T* outp = (T*)(pass == 0 ? T_out : h_out);
    const T* Sp = (const T*)(pass == 0 ? S_out : gt_out);
    if (!outp || w <= 0) continue;
    if (pass == 0) {
      HPS_CHECK_HIP(h, gather_launch<T>(s, "A", Tin, nc, nec, ex, ne, ex, ne, 1.0, 0, outp, ne, (long)ne * ne, nm));
    } else {
      static const std::string nmh = std::string("merge_gather_h<") + tname<T>() + ">";
      prof::Scope ps(s, nmh.c_str(), 0.0, 2.0 * es * (double)ne * nsrc * nm);
      hipLaunchKernelGGL((vec_gather_kernel<T>), grid_for((long)ne * nsrc, nm), dim3(256), 0, s, (const T*)h_in, nc,
                         nec, nsrc, ex, ne, 1.0, outp, nm);
      HPS_CHECK_HIP(h, hipGetLastError());
    }
    const long sO = (long)ne * w;
    const long sS = (long)ni * w;
    if (n_upd <= 0) {
      HPS_CHECK_HIP(h, gemm_call<T>(s, ne, w, ni, 1.0, Bm, ni, (long)ne * ni, Sp, w, sS, 1.0, outp, w, sO, outp, w, sO,
                                    nm));
    } else {
      HPS_CHECK_HIP(h, block_gemm<T>(s, n_upd, upd, w, 1.0, Bm, ni, (long)ne * ni, Sp, w, sS, 1.0, outp, w, sO, outp,
                                     w, sO, nm));
    }
  }
  return HPS_OK;
}

int hps_merge_level(hps_handle_t h, int dtype, int kind, int nm, int nc, int nec, int ni, int ne, int nsrc,
                    const void* T_in, const void* h_in, const int* int_rows, const int* int_cols, const int* ext,
                    const int* cols_final, const int* final_of_internal, int n_upd, const int* upd, int n_d12,
                    const int* d12_blk, int n_d21, const int* d21_blk, void* S_out, void* T_out, void* h_out,
                    void* gt_out, void* F_lu, int* perm, void* F_dinv, int* info, void* D12_out, void* D21_out,
                    void* B_out, void* F_inv_out, void* ws, size_t ws_bytes) {
  if (!h) return HPS_ERR_ARG;
  HPS_ARG(h, kind == HPS_KIND_DTN || kind == HPS_KIND_ITI, "kind must be DtN or ItI");
  HPS_ARG(h, nm > 0 && nc > 0 && nec > 0 && ni > 0 && ne > 0 && nsrc >= 0, "bad merge sizes");
  HPS_ARG(h, kind == HPS_KIND_DTN || ni % 2 == 0, "ItI merges need an even number of interface unknowns");
  HPS_ARG(h, T_in && int_rows && int_cols && ext && S_out && F_lu && perm && F_dinv && info, "null pointer");
  HPS_ARG(h, kind == HPS_KIND_DTN || final_of_internal, "ItI merges need the internal->final row map");
  HPS_ARG(h, nsrc == 0 || (h_in && gt_out), "h_in/gt_out required when nsrc > 0");
  HPS_ARG(h, n_upd >= 0 && (n_upd == 0 || upd), "bad update block list");
  for (int u = 0; u < n_upd; ++u)
    HPS_ARG(h, upd[4 * u] >= 0 && upd[4 * u + 1] > 0 && upd[4 * u] + upd[4 * u + 1] <= ne && upd[4 * u + 2] >= 0 &&
                   upd[4 * u + 3] > 0 && upd[4 * u + 2] + upd[4 * u + 3] <= ni,
            "update block out of range");
  HPS_ARG(h, n_d12 >= 0 && n_d21 >= 0 && (n_d12 == 0 || d12_blk) && (n_d21 == 0 || d21_blk), "bad coupling blocks");
  HPS_ARG(h, kind == HPS_KIND_ITI || (n_d12 == 0 && n_d21 == 0), "coupling blocks are for ItI merges");
  // every row of D12 / D21 must lie in exactly one rectangle
  for (int pass = 0; pass < 2; ++pass) {
    const int nb_ = pass == 0 ? n_d12 : n_d21;
    const int* bl = pass == 0 ? d12_blk : d21_blk;
    if (nb_ == 0) continue;
    long covered = 0;
    for (int u = 0; u < nb_; ++u) {
      const int r0 = bl[4 * u], nr = bl[4 * u + 1], k0 = bl[4 * u + 2], nk = bl[4 * u + 3];
      HPS_ARG(h, r0 >= 0 && nr > 0 && r0 + nr <= ni / 2 && k0 >= 0 && nk > 0 && k0 + nk <= ni / 2,
              "coupling block out of range");
      for (int v = 0; v < u; ++v)
        HPS_ARG(h, r0 >= bl[4 * v] + bl[4 * v + 1] || r0 + nr <= bl[4 * v], "coupling blocks overlap in rows");
      covered += nr;
    }
    HPS_ARG(h, covered == ni / 2, "coupling blocks must cover every row once");
  }
  if (ws_bytes < hps_merge_workspace(dtype, kind, nm, ni, ne, nsrc)) {
    h->err = "merge workspace too small";
    return HPS_ERR_WORKSPACE;
  }
  DISPATCH(dtype, merge_t, h, kind, nm, nc, nec, ni, ne, nsrc, T_in, h_in, int_rows, int_cols, ext, cols_final,
           final_of_internal, n_upd, upd, n_d12, d12_blk, n_d21, d21_blk, S_out, T_out, h_out, gt_out, F_lu, perm,
           F_dinv, info, D12_out, D21_out, B_out, F_inv_out, ws);
}

// ---------------------------------------------------------------------------
// up pass level (new data through stored merge factors)
// ---------------------------------------------------------------------------
size_t hps_up_workspace(int dtype, int kind, int nm, int ni, int nsrc) {
  size_t es = dtype_size(dtype);
  const int nF = kind == HPS_KIND_ITI ? ni / 2 : ni;
  // h_int, Z (ItI), and the scratch of the block-inverse solve
  return align_up((size_t)nm * ni * nsrc * es) + 2 * align_up((size_t)nm * nF * nsrc * es);
}

// the up pass solves through the block inverses when they were formed and the
// solve has few columns
static bool up_uses_fewinv(const void* F_inv, int nfac, int nsrc) {
  return F_inv != nullptr && nsrc <= 16 && fewinv_eligible(nfac);
}

template <typename T>
static int up_t(hps_handle_t h, int kind, int nm, int nc, int nec, int ni, int ne, int nsrc, const void* h_in,
                const void* inj, const int* int_rows, const int* ext, const int* final_of_internal, int n_upd,
                const int* upd, const void* F_lu, const int* perm, const void* F_dinv, const void* F_inv,
                const void* D12v, const void* D21v, const void* Bv, void* gt_out, void* h_out, void* ws) {
  hipStream_t s = h->stream;
  const int4* rows = (const int4*)int_rows;
  const int4* ex = (const int4*)ext;
  const bool iti = kind == HPS_KIND_ITI;
  const int nF = iti ? ni / 2 : ni;
  const int w = nsrc;
  const size_t es = sizeof(T);
  char* wp = (char*)ws;
  T* Yh = (T*)wp;
  wp += align_up((size_t)nm * ni * w * es);
  T* Zh = (T*)wp;
  wp += align_up((size_t)nm * nF * w * es);
  T* Ysc = (T*)wp;
  const T* Fl = (const T*)F_lu;
  const T* dinv = (const T*)F_dinv;
  const T* finv = (const T*)F_inv;
  const bool fewinv = up_uses_fewinv(F_inv, nF, w);
  const long sNF2 = (long)nF * nF;
  const long sY = (long)ni * w;
  {
    static const std::string nmh = std::string("merge_gather_h<") + tname<T>() + ">";
    prof::Scope ps(s, nmh.c_str(), 0.0, 2.0 * es * (double)ni * w * nm);
    hipLaunchKernelGGL((vec_gather_kernel<T>), grid_for((long)ni * w, nm), dim3(256), 0, s, (const T*)h_in, nc, nec, w,
                       rows, ni, -1.0, Yh, nm);
    HPS_CHECK_HIP(h, hipGetLastError());
  }
  // extra interface data (internal equation order) enters with h_int
  if (inj) HPS_CHECK_HIP(h, add2d<T>(ni, w, -1.0, (const T*)inj, w, sY, Yh, w, sY, nm, s));
  T* out = (T*)gt_out;
  if (!iti) {
    if (fewinv)
      HPS_CHECK_HIP(h, getrs_perm_fewinv<T>(ni, w, Fl, ni, (long)ni * ni, perm, finv, Yh, w, sY, out, w, sY, Ysc, nm,
                                            s));
    else
      HPS_CHECK_HIP(h, getrs_perm<T>(ni, w, Fl, ni, (long)ni * ni, perm, dinv, Yh, w, sY, out, w, sY, nm, s));
  } else {
    const long sZ = (long)nF * w;
    HPS_CHECK_HIP(h, gemm_call<T>(s, nF, w, nF, -1.0, (const T*)D12v, nF, sNF2, Yh + (long)nF * w, w, sY, 1.0, Yh, w,
                                  sY, Zh, w, sZ, nm));
    if (fewinv)
      HPS_CHECK_HIP(h, getrs_perm_fewinv<T>(nF, w, Fl, nF, sNF2, perm, finv, Zh, w, sZ, Yh, w, sY, Ysc, nm, s));
    else
      HPS_CHECK_HIP(h, getrs_perm<T>(nF, w, Fl, nF, sNF2, perm, dinv, Zh, w, sZ, Yh, w, sY, nm, s));
    HPS_CHECK_HIP(h, gemm_call<T>(s, nF, w, nF, -1.0, (const T*)D21v, nF, sNF2, Yh, w, sY, 1.0, Yh + (long)nF * w, w,
                                  sY, Yh + (long)nF * w, w, sY, nm));
    hipLaunchKernelGGL((row_scatter_kernel<T>), grid_for((long)ni * w, nm), dim3(256), 0, s, Yh, out,
                       final_of_internal, ni, w, nm);
    HPS_CHECK_HIP(h, hipGetLastError());
  }
  if (h_out) {
    T* ho = (T*)h_out;
    const long sO = (long)ne * w;
    hipLaunchKernelGGL((vec_gather_kernel<T>), grid_for((long)ne * w, nm), dim3(256), 0, s, (const T*)h_in, nc, nec, w,
                       ex, ne, 1.0, ho, nm);
    HPS_CHECK_HIP(h, hipGetLastError());
    const T* Bm = (const T*)Bv;
    if (n_upd <= 0) {
      HPS_CHECK_HIP(h, gemm_call<T>(s, ne, w, ni, 1.0, Bm, ni, (long)ne * ni, out, w, sY, 1.0, ho, w, sO, ho, w, sO, nm));
    } else {
      HPS_CHECK_HIP(h, block_gemm<T>(s, n_upd, upd, w, 1.0, Bm, ni, (long)ne * ni, out, w, sY, 1.0, ho, w, sO, ho, w,
                                     sO, nm));
    }
  }
  return HPS_OK;
}

int hps_up_level(hps_handle_t h, int dtype, int kind, int nm, int nc, int nec, int ni, int ne, int nsrc,
                 const void* h_in, const void* inj, const int* int_rows, const int* ext, const int* final_of_internal,
                 int n_upd, const int* upd, const void* F_lu, const int* perm, const void* F_dinv,
                 const void* F_inv, const void* D12, const void* D21, const void* B, void* gt_out, void* h_out,
                 void* ws, size_t ws_bytes) {
  if (!h) return HPS_ERR_ARG;
  HPS_ARG(h, kind == HPS_KIND_DTN || kind == HPS_KIND_ITI, "kind must be DtN or ItI");
  HPS_ARG(h, nm > 0 && nc > 0 && nec > 0 && ni > 0 && ne > 0 && nsrc > 0, "bad up-pass sizes");
  HPS_ARG(h, kind == HPS_KIND_DTN || (ni % 2 == 0 && D12 && D21 && final_of_internal), "ItI needs D12, D21, row map");
  HPS_ARG(h, h_in && int_rows && ext && F_lu && perm && F_dinv && gt_out, "null pointer");
  HPS_ARG(h, !h_out || B, "h_out needs B");
  HPS_ARG(h, n_upd >= 0 && (n_upd == 0 || upd), "bad update block list");
  if (ws_bytes < hps_up_workspace(dtype, kind, nm, ni, nsrc)) {
    h->err = "up workspace too small";
    return HPS_ERR_WORKSPACE;
  }
  DISPATCH(dtype, up_t, h, kind, nm, nc, nec, ni, ne, nsrc, h_in, inj, int_rows, ext, final_of_internal, n_upd, upd,
           F_lu, perm, F_dinv, F_inv, D12, D21, B, gt_out, h_out, ws);
}

// ---------------------------------------------------------------------------
// block inverses for the few-column up-pass solves
// ---------------------------------------------------------------------------
size_t hps_block_inverses_elems(int n) { return fewinv_elems(n); }

size_t hps_block_inverses_workspace(int dtype, int n, int nm) {
  return align_up(fewinv_workspace_elems(n, nm) * dtype_size(dtype));
}

template <typename T>
static int fewinv_t(hps_handle_t h, int n, int nm, const void* F_lu, const void* F_dinv, void* inv, void* ws) {
  HPS_CHECK_HIP(h, form_fewinv<T>(n, (const T*)F_lu, n, (long)n * n, (const T*)F_dinv, (T*)inv, (T*)ws, nm,
                                  h->stream));
  return HPS_OK;
}

int hps_block_inverses(hps_handle_t h, int dtype, int n, int nm, const void* F_lu, const void* F_dinv, void* inv,
                       void* ws, size_t ws_bytes) {
  if (!h) return HPS_ERR_ARG;
  HPS_ARG(h, fewinv_eligible(n), "block inverses need n > 256, a multiple of 32");
  HPS_ARG(h, nm > 0 && (long)nm * fewinv_blocks(n) <= 65535, "bad matrix count");
  HPS_ARG(h, F_lu && F_dinv && inv && ws, "null pointer");
  if (ws_bytes < hps_block_inverses_workspace(dtype, n, nm)) {
    h->err = "block-inverse workspace too small";
    return HPS_ERR_WORKSPACE;
  }
  DISPATCH(dtype, fewinv_t, h, n, nm, F_lu, F_dinv, inv, ws);
}

// ---------------------------------------------------------------------------
// down pass level and leaf apply
// ---------------------------------------------------------------------------
size_t hps_down_workspace(int dtype, int nm, int ni, int nsrc) {
  return align_up((size_t)nm * ni * nsrc * dtype_size(dtype));
}

template <typename T>
static int down_t(hps_handle_t h, int nm, int nc, int nec, int ni, int ne, int nsrc, const void* S, const void* gt,
                  const void* g_in, const int* child_src, void* g_out, void* ws) {
  hipStream_t s = h->stream;
  T* W = (T*)ws;
  GemmArgs<T> g;
  g.M = ni;
  g.N = nsrc;
  g.K = ne;
  g.alpha = 1.0;
  g.beta = gt ? 1.0 : 0.0;
  g.A = (const T*)S;
  g.lda = ne;
  g.sA = (long)ni * ne;
  g.B = (const T*)g_in;
  g.ldb = nsrc;
  g.sB = (long)ne * nsrc;
  g.C = (const T*)gt;
  g.ldc = nsrc;
  g.sC = (long)ni * nsrc;
  g.D = W;
  g.ldd = nsrc;
  g.sD = (long)ni * nsrc;
  g.batch = nm;
  HPS_CHECK_HIP(h, gemm<T>(g, s));
  {
    HPS_PROF("down_scatter", T, 0.0, 2.0 * sizeof(T) * (double)((double)nc * nec * nsrc) * nm);
    hipLaunchKernelGGL((down_scatter_kernel<T>), grid_for((long)nc * nec * nsrc, nm), dim3(256), 0, s, nc, nec, ne, ni,
                     nsrc, child_src, (const T*)g_in, W, (T*)g_out, nm);
  }
  
  HPS_CHECK_HIP(h, hipGetLastError());
  return HPS_OK;
}

int hps_down_level(hps_handle_t h, int dtype, int nm, int nc, int nec, int ni, int ne, int nsrc, const void* S,
                   const void* gt, const void* g_in, const int* child_src, void* g_out, void* ws, size_t ws_bytes) {
  if (!h) return HPS_ERR_ARG;
  HPS_ARG(h, nm > 0 && nc > 0 && nec > 0 && ni > 0 && ne > 0 && nsrc > 0, "bad down-pass sizes");
  HPS_ARG(h, S && g_in && child_src && g_out, "null pointer");
  if (ws_bytes < hps_down_workspace(dtype, nm, ni, nsrc)) {
    h->err = "down workspace too small";
    return HPS_ERR_WORKSPACE;
  }
  DISPATCH(dtype, down_t, h, nm, nc, nec, ni, ne, nsrc, S, gt, g_in, child_src, g_out, ws);
}

int hps_down_level_static(hps_handle_t h, int dtype, int nm, int nc, int nec, int ni, int ne, int nsrc,
                          const void* S, const void* gt, const void* g_in, const int* child_src, void* g_out) {
  if (!h) return HPS_ERR_ARG;
  HPS_ARG(h, nm > 0 && nc > 0 && nec > 0 && ni > 0 && ne > 0 && nsrc > 0, "bad down-pass sizes");
  HPS_ARG(h, S && g_in && child_src && g_out, "null pointer");
  HPS_ARG(h, dtype >= HPS_F64 && dtype <= HPS_C64, "bad dtype");
  const size_t need = (size_t)nm * ni * nsrc * dtype_size(dtype);
  if (need > HPS_STATIC_SCRATCH_BYTES) {
    h->err = "hps_down_level_static: nm * ni * nsrc elements (" + std::to_string(need) +
             " bytes) exceed the library's static scratch of " + std::to_string((size_t)HPS_STATIC_SCRATCH_BYTES) +
             " bytes; use hps_down_level with a workspace";
    return HPS_ERR_WORKSPACE;
  }
  const int dev = h->device;
  if (dev < 0 || dev >= HPS_MAX_DEVICES || !g_scratch_done[dev] || !g_scratch_ptr[dev]) {
    h->err = std::string("hps_down_level_static: no static scratch on this handle's device") +
             (dev >= 0 && dev < HPS_MAX_DEVICES && g_scratch_done[dev]
                  ? std::string(" (hipGetSymbolAddress: ") + hipGetErrorString(g_scratch_err[dev]) + ")"
                  : std::string());
    return HPS_ERR_HIP;
  }
  // every element of g_int is written by the product before the scatter reads it
  DISPATCH(dtype, down_t, h, nm, nc, nec, ni, ne, nsrc, S, gt, g_in, child_src, g_out, g_scratch_ptr[dev]);
}

template <typename T>
static int apply_t(hps_handle_t h, int nl, int n, int nq, int nsrc, const void* Y, const void* g, const void* v,
                   void* u) {
  GemmArgs<T> a;
  a.M = n;
  a.N = nsrc;
  a.K = nq;
  a.alpha = 1.0;
  a.beta = v ? 1.0 : 0.0;
  a.A = (const T*)Y;
  a.lda = nq;
  a.sA = (long)n * nq;
  a.B = (const T*)g;
  a.ldb = nsrc;
  a.sB = (long)nq * nsrc;
  a.C = (const T*)v;
  a.ldc = nsrc;
  a.sC = (long)n * nsrc;
  a.D = (T*)u;
  a.ldd = nsrc;
  a.sD = (long)n * nsrc;
  a.batch = nl;
  HPS_CHECK_HIP(h, gemm<T>(a, h->stream));
  return HPS_OK;
}

int hps_leaf_apply(hps_handle_t h, int dtype, int nl, int n, int nq, int nsrc, const void* Y, const void* g,
                   const void* v, void* u) {
  if (!h) return HPS_ERR_ARG;
  HPS_ARG(h, nl > 0 && n > 0 && nq > 0 && nsrc > 0, "bad sizes");
  HPS_ARG(h, Y && g && u, "null pointer");
  DISPATCH(dtype, apply_t, h, nl, n, nq, nsrc, Y, g, v, u);
}



// ---------------------------------------------------------------------------
// adjoint (reverse mode) of the solve through the stored factors
// ---------------------------------------------------------------------------
template <typename T>
static int gemm_ct_call(hps_handle_t h, int M, int N, int K, double alpha, const T* A, long lda, long sA, const T* B,
                        long ldb, long sB, double beta, const T* C, long ldc, long sC, T* D, long ldd, long sD,
                        int batch) {
  GemmArgs<T> g;
  g.M = M;
  g.N = N;
  g.K = K;
  g.alpha = alpha;
  g.beta = beta;
  g.A = A;
  g.lda = lda;
  g.sA = sA;
  g.transA = 1;
  g.B = B;
  g.ldb = ldb;
  g.sB = sB;
  g.C = C;
  g.ldc = ldc;
  g.sC = sC;
  g.D = D;
  g.ldd = ldd;
  g.sD = sD;
  g.batch = batch;
  HPS_CHECK_HIP(h, gemm<T>(g, h->stream));
  return HPS_OK;
}

template <typename T>
static int gemm_conjt_t(hps_handle_t h, int M, int N, int K, double alpha, const void* A, long lda, long sA,
                        const void* B, long ldb, long sB, double beta, const void* C, long ldc, long sC, void* D,
                        long ldd, long sD, int batch) {
  return gemm_ct_call<T>(h, M, N, K, alpha, (const T*)A, lda, sA, (const T*)B, ldb, sB, beta, (const T*)C, ldc, sC,
                         (T*)D, ldd, sD, batch);
}

int hps_gemm_batched_conjt(hps_handle_t h, int dtype, int M, int N, int K, double alpha, const void* A, long lda,
                           long sA, const void* B, long ldb, long sB, double beta, const void* C, long ldc, long sC,
                           void* D, long ldd, long sD, int batch) {
  if (!h) return HPS_ERR_ARG;
  HPS_ARG(h, M >= 0 && N >= 0 && K >= 0 && batch >= 0, "negative size");
  HPS_ARG(h, D != nullptr && (K == 0 || (A && B)), "null matrix");
  HPS_ARG(h, beta == 0.0 || C != nullptr, "beta != 0 needs C");
  DISPATCH(dtype, gemm_conjt_t, h, M, N, K, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC, D, ldd, sD, batch);
}

size_t hps_getrs_conjt_workspace(int dtype, int n, int nrhs, int batch) {
  return align_up((size_t)batch * n * nrhs * dtype_size(dtype));
}

template <typename T>
static int getrs_conjt_t(hps_handle_t h, int n, int nrhs, const void* LU, long lda, long sA, const int* perm,
                         const void* dinv, const void* R, long ldr, long sR, void* X, long ldx, long sX, int batch,
                         void* ws) {
  HPS_CHECK_HIP(h, getrs_conjt<T>(n, nrhs, (const T*)LU, lda, sA, perm, (const T*)dinv, (const T*)R, ldr, sR, (T*)X,
                                  ldx, sX, (T*)ws, batch, h->stream));
  return HPS_OK;
}

int hps_getrs_batched_conjt(hps_handle_t h, int dtype, int n, int nrhs, const void* LU, long lda, long sA,
                            const int* perm, const void* dinv, const void* R, long ldr, long sR, void* X, long ldx,
                            long sX, int batch, void* ws, size_t ws_bytes) {
  if (!h) return HPS_ERR_ARG;
  HPS_ARG(h, n >= 0 && nrhs >= 0 && batch >= 0 && lda >= n, "bad size");
  HPS_ARG(h, LU && perm && dinv && R && X && ws, "null pointer");
  HPS_ARG(h, R != X, "the solve is out of place: R and X must differ");
  if (ws_bytes < hps_getrs_conjt_workspace(dtype, n, nrhs, batch)) {
    h->err = "conjugate-transposed solve workspace too small";
    return HPS_ERR_WORKSPACE;
  }
  DISPATCH(dtype, getrs_conjt_t, h, n, nrhs, LU, lda, sA, perm, dinv, R, ldr, sR, X, ldx, sX, batch, ws);
}

template <typename T>
static int adj_down_t(hps_handle_t h, int nm, int nc, int nec, int ni, int ne, int nsrc, const void* S,
                      const void* gbar_in, const int* ext, const int* cols_final, void* gbar_ext, void* gtbar) {
  hipStream_t s = h->stream;
  const int w = nsrc;
  {
    HPS_PROF("adjoint_gather", T, 0.0, sizeof(T) * 2.0 * (double)(ni + ne) * w * nm);
    hipLaunchKernelGGL((vec_gather_kernel<T>), grid_for((long)ni * w, nm), dim3(256), 0, s, (const T*)gbar_in, nc,
                       nec, w, (const int4*)cols_final, ni, 1.0, (T*)gtbar, nm);
    hipLaunchKernelGGL((vec_gather_kernel<T>), grid_for((long)ne * w, nm), dim3(256), 0, s, (const T*)gbar_in, nc,
                       nec, w, (const int4*)ext, ne, 1.0, (T*)gbar_ext, nm);
    HPS_CHECK_HIP(h, hipGetLastError());
  }
  return gemm_ct_call<T>(h, ne, w, ni, 1.0, (const T*)S, ne, (long)ni * ne, (const T*)gtbar, w, (long)ni * w, 1.0,
                         (const T*)gbar_ext, w, (long)ne * w, (T*)gbar_ext, w, (long)ne * w, nm);
}

int hps_adjoint_down_level(hps_handle_t h, int dtype, int nm, int nc, int nec, int ni, int ne, int nsrc, const void* S,
                           const void* gbar_in, const int* ext, const int* cols_final, void* gbar_ext_out,
                           void* gtbar_out) {
  if (!h) return HPS_ERR_ARG;
  HPS_ARG(h, nm > 0 && nc > 0 && nec > 0 && ni > 0 && ne > 0 && nsrc > 0, "bad adjoint down-pass sizes");
  HPS_ARG(h, S && gbar_in && ext && cols_final && gbar_ext_out && gtbar_out, "null pointer");
  DISPATCH(dtype, adj_down_t, h, nm, nc, nec, ni, ne, nsrc, S, gbar_in, ext, cols_final, gbar_ext_out, gtbar_out);
}

size_t hps_adjoint_up_workspace(int dtype, int kind, int nm, int ni, int nsrc) {
  const size_t es = dtype_size(dtype);
  const int nf = kind == HPS_KIND_ITI ? ni / 2 : ni;
  const size_t v = align_up((size_t)nm * ni * nsrc * es);
  return 3 * v + hps_getrs_conjt_workspace(dtype, nf, nsrc, nm);
}

template <typename T>
static int adj_up_t(hps_handle_t h, int kind, int nm, int nc, int nec, int ni, int ne, int nsrc, const void* gtbar,
                    const void* hbar_out, const void* Bv, const void* F_lu, const int* perm, const void* F_dinv,
                    const void* D12v, const void* D21v, const int* final_of_internal, const int* adj_src,
                    void* hbar_children, void* ws) {
  hipStream_t s = h->stream;
  const int w = nsrc;
  const bool iti = kind == HPS_KIND_ITI;
  const int nf = iti ? ni / 2 : ni;
  const long sY = (long)ni * w;
  char* wp = (char*)ws;
  const size_t v = align_up((size_t)nm * ni * w * sizeof(T));
  T* yf = (T*)wp;
  T* y = (T*)(wp + v);
  T* x = (T*)(wp + 2 * v);
  T* W = (T*)(wp + 3 * v);
  // y (final order) = gtbar + B^H hbar_out
  HPS_CHECK_HIP(h, copy2d<T>(ni, w, 1.0, (const T*)gtbar, w, sY, yf, w, sY, nm, s));
  if (hbar_out) {
    int rc = gemm_ct_call<T>(h, ni, w, ne, 1.0, (const T*)Bv, ni, (long)ne * ni, (const T*)hbar_out, w, (long)ne * w,
                             1.0, yf, w, sY, yf, w, sY, nm);
    if (rc != HPS_OK) return rc;
  }
  hipLaunchKernelGGL((row_gather_kernel<T>), grid_for((long)ni * w, nm), dim3(256), 0, s, (const T*)yf, y,
                     final_of_internal, ni, w, nm);
  HPS_CHECK_HIP(h, hipGetLastError());
  const T* Fl = (const T*)F_lu;
  const T* dv = (const T*)F_dinv;
  const long sF = (long)nf * nf;
  if (!iti) {
    HPS_CHECK_HIP(h, getrs_conjt<T>(ni, w, Fl, ni, sF, perm, dv, y, w, sY, x, w, sY, W, nm, s));
  } else {
    // D^H = I + [[0, D21^H], [D12^H, 0]]:  x_F = W^{-H} (y_F - D21^H y_G),  x_G = y_G - D12^H x_F
    T* yF = y;
    T* yG = y + (long)nf * w;
    int rc = gemm_ct_call<T>(h, nf, w, nf, -1.0, (const T*)D21v, nf, sF, yG, w, sY, 1.0, yF, w, sY, yF, w, sY, nm);
    if (rc != HPS_OK) return rc;
    HPS_CHECK_HIP(h, getrs_conjt<T>(nf, w, Fl, nf, sF, perm, dv, yF, w, sY, x, w, sY, W, nm, s));
    rc = gemm_ct_call<T>(h, nf, w, nf, -1.0, (const T*)D12v, nf, sF, x, w, sY, 1.0, yG, w, sY, x + (long)nf * w, w, sY,
                         nm);
    if (rc != HPS_OK) return rc;
  }
  // hbar_int = -x; children's slots take hbar_int (the rows they feed) or hbar_out (exterior)
  HPS_CHECK_HIP(h, copy2d<T>(ni, w, -1.0, x, w, sY, y, w, sY, nm, s));
  {
    HPS_PROF("adjoint_scatter", T, 0.0, 2.0 * sizeof(T) * (double)nc * nec * w * nm);
    hipLaunchKernelGGL((down_scatter_kernel<T>), grid_for((long)nc * nec * w, nm), dim3(256), 0, s, nc, nec, ne, ni,
                       w, adj_src, (const T*)hbar_out, (const T*)y, (T*)hbar_children, nm);
    HPS_CHECK_HIP(h, hipGetLastError());
  }
  return HPS_OK;
}

int hps_adjoint_up_level(hps_handle_t h, int dtype, int kind, int nm, int nc, int nec, int ni, int ne, int nsrc,
                         const void* gtbar, const void* hbar_out, const void* B, const void* F_lu, const int* perm,
                         const void* F_dinv, const void* D12, const void* D21, const int* final_of_internal,
                         const int* adj_src, void* hbar_children, void* ws, size_t ws_bytes) {
  if (!h) return HPS_ERR_ARG;
  HPS_ARG(h, kind == HPS_KIND_DTN || kind == HPS_KIND_ITI, "kind must be DtN or ItI");
  HPS_ARG(h, nm > 0 && nc > 0 && nec > 0 && ni > 0 && ne > 0 && nsrc > 0, "bad adjoint up-pass sizes");
  HPS_ARG(h, kind == HPS_KIND_DTN || (ni % 2 == 0 && D12 && D21), "ItI needs D12 and D21");
  HPS_ARG(h, gtbar && F_lu && perm && F_dinv && final_of_internal && adj_src && hbar_children && ws, "null pointer");
  HPS_ARG(h, !hbar_out || B, "hbar_out needs B");
  if (ws_bytes < hps_adjoint_up_workspace(dtype, kind, nm, ni, nsrc)) {
    h->err = "adjoint up workspace too small";
    return HPS_ERR_WORKSPACE;
  }
  DISPATCH(dtype, adj_up_t, h, kind, nm, nc, nec, ni, ne, nsrc, gtbar, hbar_out, B, F_lu, perm, F_dinv, D12, D21,
           final_of_internal, adj_src, hbar_children, ws);
}

template <typename T>
static int coef_grad_t(hps_handle_t h, int nl, int n, int row0, int nsrc, int ncoef, CoefOps co, const void* R,
                       int nrow_R, const void* u, const void* lam, double* grad) {
  hipStream_t s = h->stream;
  HPS_PROF("coef_gradient", T, 0.0, sizeof(T) * (double)nl * n * nsrc * (ncoef + 2) + 8.0 * ncoef * nl * n);
  hipLaunchKernelGGL((coef_grad_kernel<T>), grid_for((long)nl * n, 1), dim3(256), 0, s, nl, n, row0, nsrc, ncoef, co,
                     (const T*)R, nrow_R, (const T*)u, (const T*)lam, grad);
  HPS_CHECK_HIP(h, hipGetLastError());
  return HPS_OK;
}

int hps_coef_gradient(hps_handle_t h, int dtype, int nl, int n, int row0, int nsrc, int nops, int ncoef,
                      const int* coef_ops, const void* R, int nrow_R, const void* u, const void* lam, double* grad) {
  if (!h) return HPS_ERR_ARG;
  HPS_ARG(h, nl > 0 && n > 0 && row0 >= 0 && row0 <= n && nsrc > 0 && nops >= 0, "bad gradient sizes");
  HPS_ARG(h, ncoef >= 1 && ncoef <= 16 && coef_ops, "ncoef must be in [1, 16] with coef_ops given");
  HPS_ARG(h, nrow_R >= nops * n, "nrow_R too small");
  HPS_ARG(h, (nops == 0 || R) && u && lam && grad, "null pointer");
  CoefOps co;
  for (int k = 0; k < ncoef; ++k) {
    HPS_ARG(h, coef_ops[k] >= -1 && coef_ops[k] < nops, "coef_ops entry out of range");
    co.op[k] = coef_ops[k];
  }
  DISPATCH(dtype, coef_grad_t, h, nl, n, row0, nsrc, ncoef, co, R, nrow_R, u, lam, grad);
}
