// Iterative refinement support: FP64 residuals of the global HPS collocation
// system for an iterate, the FP64 update x += dx with its norms, and dtype
// conversion.
//
// The uniform HPS build+solve computes, in exact arithmetic, the solution of
// one global linear system made of every leaf's collocation rows and the
// interface coupling on the leaves' Gauss boundary nodes.  A mixed-precision
// solve factors that system in FP32 / complex64 and wins FP64 accuracy back by
// refinement: r = b - K u in FP64 (these kernels), K^{-1} r through the stored
// low-precision factors (the up and down passes), u += correction.
//
// Inputs common to both residuals (per leaf l, source s):
//   R      (nl, nrow_R, w)  operator products of the iterate, one GEMM:
//          rows [op_0 u | ... | op_{nops-1} u | extra], op_k the stacked
//          differentiation matrices the coefficients use; extra is
//          [G u (nb rows) | QH u (nq rows)] for ItI and [Q u (nq rows)] for DtN
//   coeffs (ncoef, nl, n)   coefficient fields; coef_ops[k] = op row block or -1 (identity)
//   nbr    (nl, nq)         Gauss boundary slot -> neighbour slot l'*nq + j' (>= 0)
//                           or -(root boundary index + 1)
#include "hps_capi_internal.h"

namespace hps {

template <typename T>
__device__ inline double block_sum(double v, double* red) {
  const int t = threadIdx.x;
  // wave reduction, then across the block's waves
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  if ((t & 63) == 0) red[t >> 6] = v;
  __syncthreads();
  double out = 0.0;
  if (t == 0) {
    for (int k = 0; k < (int)(blockDim.x + 63) / 64; ++k) out += red[k];
  }
  return out;
}

// ItI rows (local_solve/_uniform_2D_ItI.py:120-193 builds B = [G; A_int] with
// right-hand side [P t_in; f_int]):
//   boundary rows i < nb: r_i = (P t_in)_i - (G u)_i, t_in = incoming impedance
//     on the leaf's Gauss nodes: the root data there, or minus the
//     neighbour's outgoing impedance QH u across an interior interface;
//   interior rows:        r_i = f_i - (A u)_i.
template <typename T, typename To>
__global__ __launch_bounds__(256) void residual_iti_kernel(int nl, int n, int nb, int nq, int w, int nops, int ncoef,
                                                           CoefOps ops, const T* __restrict__ coeffs,
                                                           const T* __restrict__ R, int nrow_R,
                                                           const T* __restrict__ u, const T* __restrict__ f,
                                                           const T* __restrict__ P, const T* __restrict__ g_root,
                                                           const int* __restrict__ nbr, To* __restrict__ res,
                                                           double* __restrict__ part) {
  extern __shared__ char smem_raw[];
  T* tin = (T*)smem_raw;  // (nq, w)
  __shared__ double red[8];
  const int l = blockIdx.x;
  const long rl = (long)l * nrow_R * w;
  const int qh0 = nops * n + nb;  // first QH u row
  for (int idx = threadIdx.x; idx < nq * w; idx += blockDim.x) {
    const int j = idx / w, s = idx - j * w;
    const int src = nbr[(long)l * nq + j];
    T v;
    if (src >= 0) {
      const int l2 = src / nq, j2 = src - l2 * nq;
      v = s_neg(R[((long)l2 * nrow_R + qh0 + j2) * w + s]);
    } else {
      v = g_root[(long)(-src - 1) * w + s];
    }
    tin[idx] = v;
  }
  __syncthreads();
  double acc2 = 0.0;
  for (int idx = threadIdx.x; idx < n * w; idx += blockDim.x) {
    const int i = idx / w, s = idx - i * w;
    T r;
    if (i < nb) {
      T a = s_zero<T>();
      const T* Pi = P + (long)i * nq;
      for (int j = 0; j < nq; ++j) a = s_add(a, s_mul(Pi[j], tin[j * w + s]));
      r = s_sub(a, R[rl + (long)(nops * n + i) * w + s]);
    } else {
      T a = s_zero<T>();
      for (int k = 0; k < ncoef; ++k) {
        const T c = coeffs[((long)k * nl + l) * n + i];
        const int op = ops.op[k];
        const T x = op < 0 ? u[((long)l * n + i) * w + s] : R[rl + (long)(op * n + i) * w + s];
        a = s_add(a, s_mul(c, x));
      }
      const T fi = f ? f[((long)l * n + i) * w + s] : s_zero<T>();
      r = s_sub(fi, a);
    }
    res[((long)l * n + i) * w + s] = s_cvt<To>(r);
    acc2 += s_norm2(r);
  }
  const double tot = block_sum<T>(acc2, red);
  if (threadIdx.x == 0) part[l] = tot;
}

// DtN rows (merge/_schur_complement.py:183-237 eliminates exactly these):
//   interior rows: r_i = f_i - (A u)_i with u's boundary rows = P g_leaf;
//   interface flux balance on every interior Gauss node, reported once (on the
//     slot with the smaller global index): mis = (Q u_a)(node) + (Q u_b)(node);
//   root boundary: g_root - g_leaf.
template <typename T, typename To>
__global__ __launch_bounds__(256) void residual_dtn_kernel(int nl, int n, int nb, int nq, int w, int nops, int ncoef,
                                                           CoefOps ops, const T* __restrict__ coeffs,
                                                           const T* __restrict__ R, int nrow_R,
                                                           const T* __restrict__ u, const T* __restrict__ f,
                                                           const T* __restrict__ g_leaf,
                                                           const T* __restrict__ g_root,
                                                           const int* __restrict__ nbr, To* __restrict__ res,
                                                           To* __restrict__ hinj, To* __restrict__ groot_res,
                                                           double* __restrict__ part) {
  __shared__ double red[8];
  const int l = blockIdx.x;
  const long rl = (long)l * nrow_R * w;
  const int q0 = nops * n;  // first Q u row
  double acc2 = 0.0;
  for (int idx = threadIdx.x; idx < n * w; idx += blockDim.x) {
    const int i = idx / w, s = idx - i * w;
    T r = s_zero<T>();
    if (i >= nb) {
      T a = s_zero<T>();
      for (int k = 0; k < ncoef; ++k) {
        const T c = coeffs[((long)k * nl + l) * n + i];
        const int op = ops.op[k];
        const T x = op < 0 ? u[((long)l * n + i) * w + s] : R[rl + (long)(op * n + i) * w + s];
        a = s_add(a, s_mul(c, x));
      }
      const T fi = f ? f[((long)l * n + i) * w + s] : s_zero<T>();
      r = s_sub(fi, a);
    }
    res[((long)l * n + i) * w + s] = s_cvt<To>(r);
    acc2 += s_norm2(r);
  }
  for (int idx = threadIdx.x; idx < nq * w; idx += blockDim.x) {
    const int j = idx / w, s = idx - j * w;
    const long me = (long)l * nq + j;
    const int src = nbr[me];
    T x = s_zero<T>();
    if (src >= 0) {
      if (me < src) {
        const int l2 = src / nq, j2 = src - l2 * nq;
        x = s_add(R[rl + (long)(q0 + j) * w + s], R[((long)l2 * nrow_R + q0 + j2) * w + s]);
      }
    } else {
      const long ri = (long)(-src - 1) * w + s;
      const T gr = s_sub(g_root[ri], g_leaf[me * w + s]);
      groot_res[ri] = s_cvt<To>(gr);
      acc2 += s_norm2(gr);
    }
    hinj[me * w + s] = s_cvt<To>(x);
    acc2 += s_norm2(x);
  }
  const double tot = block_sum<T>(acc2, red);
  if (threadIdx.x == 0) part[l] = tot;
}

// x += dx (dx possibly in lower precision); part[2b] = sum |dx|^2, part[2b+1] = sum |x|^2
template <typename T, typename Td>
__global__ __launch_bounds__(256) void update_kernel(long n, T* __restrict__ x, const Td* __restrict__ dx,
                                                     double* __restrict__ part) {
  __shared__ double red[8];
  double ad = 0.0, ax = 0.0;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const T d = s_cvt<T>(dx[i]);
    const T v = s_add(x[i], d);
    x[i] = v;
    ad += s_norm2(d);
    ax += s_norm2(v);
  }
  const double td = block_sum<T>(ad, red);
  __syncthreads();
  const double tx = block_sum<T>(ax, red);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = td;
    part[2 * blockIdx.x + 1] = tx;
  }
}

template <typename To, typename Ti>
__global__ __launch_bounds__(256) void convert_kernel(long n, const Ti* __restrict__ src, To* __restrict__ dst) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    dst[i] = s_cvt<To>(src[i]);
}

}  // namespace hps

// ===========================================================================
// C ABI
// ===========================================================================
static bool fill_ops(hps_handle_t h, int ncoef, const int* coef_ops, int nops, CoefOps& co) {
  if (ncoef < 1 || ncoef > 16 || !coef_ops) {
    h->err = "ncoef must be in [1, 16] with coef_ops given";
    return false;
  }
  for (int k = 0; k < ncoef; ++k) {
    if (coef_ops[k] < -1 || coef_ops[k] >= nops) {
      h->err = "coef_ops entry out of range";
      return false;
    }
    co.op[k] = coef_ops[k];
  }
  return true;
}

template <typename T, typename To>
static int residual_iti_t(hps_handle_t h, int nl, int n, int nb, int nq, int w, int nops, int ncoef, CoefOps co,
                          const void* coeffs, const void* R, int nrow_R, const void* u, const void* f, const void* P,
                          const void* g_root, const int* nbr, void* res, double* part) {
  hipStream_t s = h->stream;
  HPS_PROF("residual_iti", T, 0.0,
           sizeof(T) * ((double)nl * n * w * 3 + (double)nl * n * ncoef + (double)nl * nq * w) +
               sizeof(To) * (double)nl * n * w);
  hipLaunchKernelGGL((residual_iti_kernel<T, To>), dim3(nl), dim3(256), sizeof(T) * nq * w, s, nl, n, nb, nq, w, nops,
                     ncoef, co, (const T*)coeffs, (const T*)R, nrow_R, (const T*)u, (const T*)f, (const T*)P,
                     (const T*)g_root, nbr, (To*)res, part);
  HPS_CHECK_HIP(h, hipGetLastError());
  return HPS_OK;
}

template <typename T, typename To>
static int residual_dtn_t(hps_handle_t h, int nl, int n, int nb, int nq, int w, int nops, int ncoef, CoefOps co,
                          const void* coeffs, const void* R, int nrow_R, const void* u, const void* f,
                          const void* g_leaf, const void* g_root, const int* nbr, void* res, void* hinj,
                          void* groot_res, double* part) {
  hipStream_t s = h->stream;
  HPS_PROF("residual_dtn", T, 0.0,
           sizeof(T) * ((double)nl * n * w * 3 + (double)nl * n * ncoef + 3.0 * nl * nq * w) +
               sizeof(To) * ((double)nl * (n + nq) * w));
  hipLaunchKernelGGL((residual_dtn_kernel<T, To>), dim3(nl), dim3(256), 0, s, nl, n, nb, nq, w, nops, ncoef, co,
                     (const T*)coeffs, (const T*)R, nrow_R, (const T*)u, (const T*)f, (const T*)g_leaf,
                     (const T*)g_root, nbr, (To*)res, (To*)hinj, (To*)groot_res, part);
  HPS_CHECK_HIP(h, hipGetLastError());
  return HPS_OK;
}

// output precision: the working type itself or its single-precision partner
static int lowered(int dtype) { return dtype == HPS_F64 ? HPS_F32 : HPS_C64; }

int hps_residual_iti(hps_handle_t h, int dtype, int out_dtype, int nl, int n, int nb, int nq, int nsrc, int nops,
                     int ncoef, const int* coef_ops, const void* coeffs, const void* R, int nrow_R, const void* u,
                     const void* f, const void* P, const void* g_root, const int* nbr, void* res, double* part) {
  if (!h) return HPS_ERR_ARG;
  HPS_ARG(h, dtype == HPS_C128, "ItI residuals are complex128");
  HPS_ARG(h, out_dtype == dtype || out_dtype == lowered(dtype), "out_dtype must be dtype or its FP32 partner");
  HPS_ARG(h, nl > 0 && n > 0 && nb > 0 && nb < n && nq > 0 && nsrc > 0 && nops >= 0, "bad residual sizes");
  HPS_ARG(h, nrow_R == nops * n + nb + nq, "nrow_R must be nops*n + nb + nq");
  HPS_ARG(h, (long)nq * nsrc * 16 <= 65536, "nq * nsrc too large for the LDS stage");
  HPS_ARG(h, coeffs && R && u && P && g_root && nbr && res && part, "null pointer");
  CoefOps co;
  if (!fill_ops(h, ncoef, coef_ops, nops, co)) return HPS_ERR_ARG;
  if (out_dtype == HPS_C128)
    return residual_iti_t<c128, c128>(h, nl, n, nb, nq, nsrc, nops, ncoef, co, coeffs, R, nrow_R, u, f, P, g_root,
                                      nbr, res, part);
  return residual_iti_t<c128, c64>(h, nl, n, nb, nq, nsrc, nops, ncoef, co, coeffs, R, nrow_R, u, f, P, g_root, nbr,
                                   res, part);
}

int hps_residual_dtn(hps_handle_t h, int dtype, int out_dtype, int nl, int n, int nb, int nq, int nsrc, int nops,
                     int ncoef, const int* coef_ops, const void* coeffs, const void* R, int nrow_R, const void* u,
                     const void* f, const void* g_leaf, const void* g_root, const int* nbr, void* res, void* hinj,
                     void* groot_res, double* part) {
  if (!h) return HPS_ERR_ARG;
  HPS_ARG(h, dtype == HPS_F64 || dtype == HPS_C128, "DtN residuals are float64 or complex128");
  HPS_ARG(h, out_dtype == dtype || out_dtype == lowered(dtype), "out_dtype must be dtype or its FP32 partner");
  HPS_ARG(h, nl > 0 && n > 0 && nb > 0 && nb < n && nq > 0 && nsrc > 0 && nops >= 0, "bad residual sizes");
  HPS_ARG(h, nrow_R == nops * n + nq, "nrow_R must be nops*n + nq");
  HPS_ARG(h, coeffs && R && u && g_leaf && g_root && nbr && res && hinj && groot_res && part, "null pointer");
  CoefOps co;
  if (!fill_ops(h, ncoef, coef_ops, nops, co)) return HPS_ERR_ARG;
#define RD(T, To)                                                                                                  \
  return residual_dtn_t<T, To>(h, nl, n, nb, nq, nsrc, nops, ncoef, co, coeffs, R, nrow_R, u, f, g_leaf, g_root, \
                               nbr, res, hinj, groot_res, part)
  if (dtype == HPS_F64) {
    if (out_dtype == HPS_F64) RD(double, double);
    RD(double, float);
  }
  if (out_dtype == HPS_C128) RD(c128, c128);
  RD(c128, c64);
#undef RD
}

template <typename T, typename Td>
static int update_t(hps_handle_t h, long n, void* x, const void* dx, int nparts, double* part) {
  hipStream_t s = h->stream;
  HPS_PROF("refine_update", T, 0.0, (2.0 * sizeof(T) + sizeof(Td)) * (double)n);
  hipLaunchKernelGGL((update_kernel<T, Td>), dim3(nparts), dim3(256), 0, s, n, (T*)x, (const Td*)dx, part);
  HPS_CHECK_HIP(h, hipGetLastError());
  return HPS_OK;
}

int hps_refine_update(hps_handle_t h, int dtype, int delta_dtype, long n, void* x, const void* dx, int nparts,
                      double* part) {
  if (!h) return HPS_ERR_ARG;
  HPS_ARG(h, dtype == HPS_F64 || dtype == HPS_C128, "the iterate is float64 or complex128");
  HPS_ARG(h, delta_dtype == dtype || delta_dtype == lowered(dtype), "delta_dtype must be dtype or its FP32 partner");
  HPS_ARG(h, n >= 0 && nparts > 0 && nparts <= 65535, "bad update sizes");
  HPS_ARG(h, (n == 0 || (x && dx)) && part, "null pointer");
  if (dtype == HPS_F64) {
    if (delta_dtype == HPS_F64) return update_t<double, double>(h, n, x, dx, nparts, part);
    return update_t<double, float>(h, n, x, dx, nparts, part);
  }
  if (delta_dtype == HPS_C128) return update_t<c128, c128>(h, n, x, dx, nparts, part);
  return update_t<c128, c64>(h, n, x, dx, nparts, part);
}

template <typename To, typename Ti>
static int convert_t(hps_handle_t h, long n, const void* src, void* dst) {
  hipStream_t s = h->stream;
  if (n == 0) return HPS_OK;
  HPS_PROF("convert", To, 0.0, (double)(sizeof(To) + sizeof(Ti)) * n);
  long blocks = (n + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL((convert_kernel<To, Ti>), dim3((unsigned)blocks), dim3(256), 0, s, n, (const Ti*)src, (To*)dst);
  HPS_CHECK_HIP(h, hipGetLastError());
  return HPS_OK;
}

template <typename To>
static int convert_from(hps_handle_t h, int src_dtype, long n, const void* src, void* dst) {
  switch (src_dtype) {
    case HPS_F64: return convert_t<To, double>(h, n, src, dst);
    case HPS_F32: return convert_t<To, float>(h, n, src, dst);
    case HPS_C128: return convert_t<To, c128>(h, n, src, dst);
    case HPS_C64: return convert_t<To, c64>(h, n, src, dst);
  }
  h->err = "bad source dtype";
  return HPS_ERR_ARG;
}

int hps_convert(hps_handle_t h, int src_dtype, int dst_dtype, long n, const void* src, void* dst) {
  if (!h) return HPS_ERR_ARG;
  HPS_ARG(h, n >= 0, "negative size");
  HPS_ARG(h, n == 0 || (src && dst && src != dst), "null or aliased pointer");
  switch (dst_dtype) {
    case HPS_F64: return convert_from<double>(h, src_dtype, n, src, dst);
    case HPS_F32: return convert_from<float>(h, src_dtype, n, src, dst);
    case HPS_C128: return convert_from<c128>(h, src_dtype, n, src, dst);
    case HPS_C64: return convert_from<c64>(h, src_dtype, n, src, dst);
  }
  h->err = "bad destination dtype";
  return HPS_ERR_ARG;
}
