// Batched GEMM for gfx950: D = alpha * A @ B + beta * C.
//
// * gemm_mfma_kernel: 64x64x16 block tile, 256 threads = 4 waves of 64 in a
//   2x2 arrangement, each wave owning a 32x32 sub-tile as 2x2 MFMA 16x16x4
//   blocks (v_mfma_f64_16x16x4f64 / v_mfma_f32_16x16x4f32).  Complex types
//   run 4 real MFMAs per k-step on planar (re/im split) LDS tiles.
// * N <= 16 (several right-hand sides): the same kernel with the four waves
//   stacked along M, one 16-wide MFMA column block.
// * gemm_skinny_kernel: N <= 2 (down-pass / leaf-apply GEMV shapes).  One
//   wave per output row, lanes stride the K dimension with coalesced loads
//   of A's row, then a wave reduction.  These shapes are HBM-bound on A.
#include <string>

#include "gemm_kernels.h"
#include "hps_prof.h"

namespace hps {

template <typename T>
hipError_t gemm(const GemmArgs<T>& a, hipStream_t s) {
  if (a.M <= 0 || a.N <= 0 || a.batch <= 0) return hipSuccess;
  // the triangular block step has one kernel: plain A, K <= 32, N >= 3
  if (a.triA != 0 && (a.transA || a.K > 32 || a.N <= 16 || a.M > 32)) return hipErrorInvalidValue;
  unsigned gy = (unsigned)(a.batch < 65535 ? a.batch : 65535);
  const double flops = 2.0 * a.M * a.N * (double)a.K * a.batch * flop_mult<T>();
  const double bytes = sizeof(T) * ((double)a.M * a.K * (a.sA ? a.batch : 1) + (double)a.K * a.N * (a.sB ? a.batch : 1) +
                                    (double)a.M * a.N * a.batch * (a.beta != 0.0 ? 2 : 1));
  if (a.transA) {
    // op(A) = A^H: the adjoint products (S^H, B^H, Y^H, D12^H ... and the
    // conjugate-transposed triangular updates)
    if (a.N <= 2) {
      static const std::string nm = std::string("gemv_conjt<") + tname<T>() + ">";
      prof::Scope ps(s, nm.c_str(), flops, bytes);
      dim3 grid(ceil_div(a.M, 256), gy);
      if (a.N <= 1)
        hipLaunchKernelGGL((gemv_conjt_kernel<T, 1>), grid, dim3(256), 0, s, a);
      else
        hipLaunchKernelGGL((gemv_conjt_kernel<T, 2>), grid, dim3(256), 0, s, a);
    } else if (a.N <= 16) {
      static const std::string nm = std::string("gemm_mfma_64x16_conjt<") + tname<T>() + ">";
      prof::Scope ps(s, nm.c_str(), flops, bytes);
      dim3 grid((unsigned)ceil_div(a.M, 64), gy);
      hipLaunchKernelGGL((gemm_mfma_kernel<T, 1, 1, 4, 4, true>), grid, dim3(256), 0, s, a, 1);
    } else {
      static const std::string nm = std::string("gemm_mfma_64x64_conjt<") + tname<T>() + ">";
      prof::Scope ps(s, nm.c_str(), flops, bytes);
      int tiles_n = ceil_div(a.N, 64), tiles_m = ceil_div(a.M, 64);
      dim3 grid((unsigned)(tiles_m * tiles_n), gy);
      hipLaunchKernelGGL((gemm_mfma_kernel<T, 2, 2, 3, 2, true>), grid, dim3(256), 0, s, a, tiles_n);
    }
    return hipGetLastError();
  }
  if (a.N <= 2) {
    static const std::string nm = std::string("gemm_skinny<") + tname<T>() + ">";
    prof::Scope ps(s, nm.c_str(), flops, bytes);
    // lanes per row: about one element of the row per lane
    const int lpr = a.K <= 64 ? 4 : (a.K <= 512 ? 16 : 64);
    dim3 grid(ceil_div(a.M, 256 / lpr), gy);
    if (lpr == 4) {
      if (a.N <= 1)
        hipLaunchKernelGGL((gemm_skinny_kernel<T, 1, 4>), grid, dim3(256), 0, s, a);
      else
        hipLaunchKernelGGL((gemm_skinny_kernel<T, 2, 4>), grid, dim3(256), 0, s, a);
    } else if (lpr == 16) {
      if (a.N <= 1)
        hipLaunchKernelGGL((gemm_skinny_kernel<T, 1, 16>), grid, dim3(256), 0, s, a);
      else
        hipLaunchKernelGGL((gemm_skinny_kernel<T, 2, 16>), grid, dim3(256), 0, s, a);
    } else if (a.N <= 1) {
      hipLaunchKernelGGL((gemm_skinny_kernel<T, 1>), grid, dim3(256), 0, s, a);
    } else {
      hipLaunchKernelGGL((gemm_skinny_kernel<T, 2>), grid, dim3(256), 0, s, a);
    }
  } else if (a.N <= 16) {
    // one 16-wide MFMA column block; the four waves stack along M (64-row tiles)
    static const std::string nm = std::string("gemm_mfma_64x16<") + tname<T>() + ">";
    prof::Scope ps(s, nm.c_str(), flops, bytes);
    int tiles_m = ceil_div(a.M, 64);
    dim3 grid((unsigned)tiles_m, gy);
    hipLaunchKernelGGL((gemm_mfma_kernel<T, 1, 1, 4, 4>), grid, dim3(256), 0, s, a, 1);
  } else {
    static const std::string nm64 = std::string("gemm_mfma_64x64<") + tname<T>() + ">";
    static const std::string nm32 = std::string("gemm_mfma_64x32<") + tname<T>() + ">";
    static const std::string nmk32 = std::string("gemm_k32<") + tname<T>() + ">";
    static const std::string nmtri = std::string("trsm_k32<") + tname<T>() + ">";
    if (a.K <= 32) {
      prof::Scope ps(s, (a.triA ? nmtri : nmk32).c_str(), flops, bytes);
      if (a.M <= 32) {
        // 32 x 64 tiles (the triangular block steps and other 32-row products)
        int tiles_n = ceil_div(a.N, 64);
        dim3 grid((unsigned)tiles_n, gy);
        if (a.triA == 1)
          hipLaunchKernelGGL((gemm_k32_kernel<T, 1, 1, 2>), grid, dim3(256), 0, s, a, tiles_n);
        else if (a.triA == 2)
          hipLaunchKernelGGL((gemm_k32_kernel<T, 2, 1, 2>), grid, dim3(256), 0, s, a, tiles_n);
        else
          hipLaunchKernelGGL((gemm_k32_kernel<T, 0, 1, 2>), grid, dim3(256), 0, s, a, tiles_n);
      } else {
        int tiles_n = ceil_div(a.N, 32), tiles_m = ceil_div(a.M, 64);
        dim3 grid((unsigned)(tiles_m * tiles_n), gy);
        hipLaunchKernelGGL((gemm_k32_kernel<T>), grid, dim3(256), 0, s, a, tiles_n);
      }
      return hipGetLastError();
    }
    const bool short_k = a.K <= 64 && a.N <= 256;
    prof::Scope ps(s, (short_k ? nm32 : nm64).c_str(), flops, bytes);
    // Tile choice measured on MI355X over the solver's shapes: 64x64 tiles
    // with 3 workgroups per CU for general shapes (2 for complex128); 64x32
    // tiles with 2 per CU for the short-K batched updates of the leaf LU and
    // solves.  128-wide tiles measured slower (one workgroup per CU).
    if (short_k) {
      int tiles_n = ceil_div(a.N, 32), tiles_m = ceil_div(a.M, 64);
      dim3 grid((unsigned)(tiles_m * tiles_n), gy);
      hipLaunchKernelGGL((gemm_mfma_kernel<T, 2, 1, 2>), grid, dim3(256), 0, s, a, tiles_n);
    } else {
      int tiles_n = ceil_div(a.N, 64), tiles_m = ceil_div(a.M, 64);
      dim3 grid((unsigned)(tiles_m * tiles_n), gy);
      if constexpr (sizeof(T) == 16)  // 3 workgroups per CU would force complex128 to spill registers
        hipLaunchKernelGGL((gemm_mfma_kernel<T, 2, 2, 2>), grid, dim3(256), 0, s, a, tiles_n);
      else
        hipLaunchKernelGGL((gemm_mfma_kernel<T, 2, 2, 3, 2, false, is_cplx<T>::value>), grid, dim3(256), 0, s, a,
                           tiles_n);
    }
  }
  return hipGetLastError();
}

template <typename T>
hipError_t gemm_grouped(const GemmArgs<T>& a, int n_blk, const int* blk, hipStream_t s) {
  if (a.N <= 0 || a.batch <= 0 || n_blk <= 0) return hipSuccess;
  if (a.transA || a.triA) return hipErrorInvalidValue;
  // one launch when every group takes the 64 x 64 tile path (N > 256, K > 32);
  // otherwise one product per group
  bool one = a.N > 256;
  for (int u = 0; u < n_blk && one; ++u) one = blk[4 * u + 3] > 32;
  // groups that write the same rows run one after another (each reads the
  // previous one's D as its C)
  for (int u = 0; u < n_blk && one; ++u)
    for (int v = u + 1; v < n_blk && one; ++v)
      one = blk[4 * u] + blk[4 * u + 1] <= blk[4 * v] || blk[4 * v] + blk[4 * v + 1] <= blk[4 * u];
  auto group_args = [&](int u) {
    GemmArgs<T> g = a;
    const int r0 = blk[4 * u], k0 = blk[4 * u + 2];
    g.M = blk[4 * u + 1];
    g.K = blk[4 * u + 3];
    g.A = a.A + (long)r0 * a.lda + k0;
    g.B = a.B + (long)k0 * a.ldb;
    g.C = a.C ? a.C + (long)r0 * a.ldc : nullptr;
    g.D = a.D + (long)r0 * a.ldd;
    return g;
  };
  if (!one) {
    for (int u = 0; u < n_blk; ++u) {
      hipError_t e = gemm<T>(group_args(u), s);
      if (e != hipSuccess) return e;
    }
    return hipSuccess;
  }
  static const std::string nm = std::string("gemm_mfma_64x64<") + tname<T>() + ">";
  const int tiles_n = ceil_div(a.N, 64);
  const unsigned gy = (unsigned)(a.batch < 65535 ? a.batch : 65535);
  for (int u0 = 0; u0 < n_blk; u0 += GEMM_MAX_GROUPS) {
    GemmGroups gr;
    gr.n = n_blk - u0 < GEMM_MAX_GROUPS ? n_blk - u0 : GEMM_MAX_GROUPS;
    double flops = 0, bytes = 0;
    for (int i = 0; i < gr.n; ++i) {
      const int u = u0 + i;
      gr.r0[i] = blk[4 * u];
      gr.nr[i] = blk[4 * u + 1];
      gr.k0[i] = blk[4 * u + 2];
      gr.nk[i] = blk[4 * u + 3];
      gr.start[i + 1] = gr.start[i] + ceil_div(gr.nr[i], 64) * tiles_n;
      flops += 2.0 * gr.nr[i] * a.N * (double)gr.nk[i] * a.batch * flop_mult<T>();
      bytes += sizeof(T) * ((double)gr.nr[i] * gr.nk[i] * (a.sA ? a.batch : 1) +
                            (double)gr.nk[i] * a.N * (a.sB ? a.batch : 1) +
                            (double)gr.nr[i] * a.N * a.batch * (a.beta != 0.0 ? 2 : 1));
    }
    if (gr.start[gr.n] == 0) continue;  // blocks without rows: nothing to launch
    prof::Scope ps(s, nm.c_str(), flops, bytes);
    dim3 grid((unsigned)gr.start[gr.n], gy);
    if constexpr (sizeof(T) == 16)
      hipLaunchKernelGGL((gemm_mfma_grouped_kernel<T, 2, 2, 2, false>), grid, dim3(256), 0, s, a, gr, tiles_n);
    else
      hipLaunchKernelGGL((gemm_mfma_grouped_kernel<T, 2, 2, 3, is_cplx<T>::value>), grid, dim3(256), 0, s, a, gr,
                         tiles_n);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

template hipError_t gemm<double>(const GemmArgs<double>&, hipStream_t);
template hipError_t gemm<float>(const GemmArgs<float>&, hipStream_t);
template hipError_t gemm<c128>(const GemmArgs<c128>&, hipStream_t);
template hipError_t gemm<c64>(const GemmArgs<c64>&, hipStream_t);
template hipError_t gemm_grouped<double>(const GemmArgs<double>&, int, const int*, hipStream_t);
template hipError_t gemm_grouped<float>(const GemmArgs<float>&, int, const int*, hipStream_t);
template hipError_t gemm_grouped<c128>(const GemmArgs<c128>&, int, const int*, hipStream_t);
template hipError_t gemm_grouped<c64>(const GemmArgs<c64>&, int, const int*, hipStream_t);

}  // namespace hps
