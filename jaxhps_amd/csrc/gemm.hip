// Batched GEMM for gfx950: D = alpha * A @ B + beta * C.
//
// * gemm_mfma_kernel: 64x64x16 block tile, 256 threads = 4 waves of 64 in a
//   2x2 arrangement, each wave owning a 32x32 sub-tile as 2x2 MFMA 16x16x4
//   blocks (v_mfma_f64_16x16x4f64 / v_mfma_f32_16x16x4f32).  Complex types
//   run 4 real MFMAs per k-step on planar (re/im split) LDS tiles.
// * gemm_skinny_kernel: N <= 16 (down-pass / leaf-apply GEMV shapes).  One
//   wave per output row, lanes stride the K dimension with coalesced loads
//   of A's row, then a wave reduction.  These shapes are HBM-bound on A.
#include <string>

#include "hps_kernels.h"
#include "hps_prof.h"

namespace hps {

typedef double dbl4 __attribute__((ext_vector_type(4)));
typedef float flt4 __attribute__((ext_vector_type(4)));

// Result layout of the 16x16x4 MFMAs (measured on gfx950 with an in-bounds
// single-wave probe): register r of lane l holds
//   f64: C[l/16 + 4r][l%16]        f32: C[4(l/16) + r][l%16]
// while lane l supplies A[l%16][l/16] and B[l/16][l%16] for both.
template <typename R> struct Mfma;
template <> struct Mfma<double> {
  using acc_t = dbl4;
  static __device__ __forceinline__ int out_row(int lk, int r) { return lk + 4 * r; }
  static __device__ __forceinline__ acc_t mma(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
};
template <> struct Mfma<float> {
  using acc_t = flt4;
  static __device__ __forceinline__ int out_row(int lk, int r) { return 4 * lk + r; }
  static __device__ __forceinline__ acc_t mma(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
};

template <typename T> __device__ __forceinline__ T make_val(typename real_of<T>::type re, typename real_of<T>::type im);
template <> __device__ __forceinline__ double make_val<double>(double re, double) { return re; }
template <> __device__ __forceinline__ float make_val<float>(float re, float) { return re; }
template <> __device__ __forceinline__ c128 make_val<c128>(double re, double im) { return c128{re, im}; }
template <> __device__ __forceinline__ c64 make_val<c64>(float re, float im) { return c64{re, im}; }

constexpr int GBM = 64, GBN = 64, GBK = 16;

template <typename T>
__global__ __launch_bounds__(256) void gemm_mfma_kernel(GemmArgs<T> p, int tiles_n) {
  using R = typename real_of<T>::type;
  using M = Mfma<R>;
  using acc_t = typename M::acc_t;
  constexpr bool CPLX = is_cplx<T>::value;
  __shared__ R As_re[GBK][GBM + 1];
  __shared__ R Bs_re[GBK][GBN + 1];
  __shared__ R As_im[CPLX ? GBK : 1][CPLX ? GBM + 1 : 1];
  __shared__ R Bs_im[CPLX ? GBK : 1][CPLX ? GBN + 1 : 1];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wr = wave >> 1, wc = wave & 1;
  const int tm = blockIdx.x / tiles_n, tn = blockIdx.x % tiles_n;
  const int m0 = tm * GBM, n0 = tn * GBN;
  const int li = lane & 15, lk = lane >> 4;

  for (int b = blockIdx.y; b < p.batch; b += gridDim.y) {
    const T* A = p.A + (long)b * p.sA;
    const T* B = p.B + (long)b * p.sB;
    acc_t acc_re[2][2], acc_im[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        acc_re[i][j] = acc_t{0, 0, 0, 0};
        acc_im[i][j] = acc_t{0, 0, 0, 0};
      }

    for (int k0 = 0; k0 < p.K; k0 += GBK) {
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        int e = tid + 256 * i;
        int r = e >> 4, c = e & 15;
        int gr = m0 + r, gc = k0 + c;
        T v = (gr < p.M && gc < p.K) ? A[(long)gr * p.lda + gc] : s_zero<T>();
        As_re[c][r] = s_re(v);
        if constexpr (CPLX) As_im[c][r] = s_im(v);
      }
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        int e = tid + 256 * i;
        int r = e >> 6, c = e & 63;
        int gr = k0 + r, gc = n0 + c;
        T v = (gr < p.K && gc < p.N) ? B[(long)gr * p.ldb + gc] : s_zero<T>();
        Bs_re[r][c] = s_re(v);
        if constexpr (CPLX) Bs_im[r][c] = s_im(v);
      }
      __syncthreads();
#pragma unroll
      for (int kk = 0; kk < GBK; kk += 4) {
        R a_re[2], b_re[2], a_im[2], b_im[2];
#pragma unroll
        for (int t = 0; t < 2; ++t) {
          a_re[t] = As_re[kk + lk][wr * 32 + t * 16 + li];
          b_re[t] = Bs_re[kk + lk][wc * 32 + t * 16 + li];
          if constexpr (CPLX) {
            a_im[t] = As_im[kk + lk][wr * 32 + t * 16 + li];
            b_im[t] = Bs_im[kk + lk][wc * 32 + t * 16 + li];
          }
        }
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
          for (int j = 0; j < 2; ++j) {
            acc_re[i][j] = M::mma(a_re[i], b_re[j], acc_re[i][j]);
            if constexpr (CPLX) {
              acc_re[i][j] = M::mma(-a_im[i], b_im[j], acc_re[i][j]);
              acc_im[i][j] = M::mma(a_re[i], b_im[j], acc_im[i][j]);
              acc_im[i][j] = M::mma(a_im[i], b_re[j], acc_im[i][j]);
            }
          }
      }
      __syncthreads();
    }

    const T* C = p.C ? p.C + (long)b * p.sC : nullptr;
    T* D = p.D + (long)b * p.sD;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          int row = m0 + wr * 32 + i * 16 + M::out_row(lk, r);
          int col = n0 + wc * 32 + j * 16 + li;
          if (row < p.M && col < p.N) {
            T v = make_val<T>(acc_re[i][j][r], CPLX ? acc_im[i][j][r] : R(0));
            v = s_scale(v, p.alpha);
            if (p.beta != 0.0) v = s_add(v, s_scale(C[(long)row * p.ldc + col], p.beta));
            D[(long)row * p.ldd + col] = v;
          }
        }
  }
}

template <typename R>
__device__ __forceinline__ R wave_sum(R v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

template <typename T, int NMAX>
__global__ __launch_bounds__(256) void gemm_skinny_kernel(GemmArgs<T> p) {
  using R = typename real_of<T>::type;
  constexpr bool CPLX = is_cplx<T>::value;
  const int lane = threadIdx.x & 63;
  const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (row >= p.M) return;
  for (int b = blockIdx.y; b < p.batch; b += gridDim.y) {
    const T* A = p.A + (long)b * p.sA + (long)row * p.lda;
    const T* B = p.B + (long)b * p.sB;
    R acc_re[NMAX], acc_im[NMAX];
#pragma unroll
    for (int n = 0; n < NMAX; ++n) {
      acc_re[n] = R(0);
      acc_im[n] = R(0);
    }
    for (int k = lane; k < p.K; k += 64) {
      T a = A[k];
      const T* Bk = B + (long)k * p.ldb;
#pragma unroll
      for (int n = 0; n < NMAX; ++n) {
        if (n < p.N) {
          T bv = Bk[n];
          if constexpr (CPLX) {
            acc_re[n] += s_re(a) * s_re(bv) - s_im(a) * s_im(bv);
            acc_im[n] += s_re(a) * s_im(bv) + s_im(a) * s_re(bv);
          } else {
            acc_re[n] += s_re(a) * s_re(bv);
          }
        }
      }
    }
#pragma unroll
    for (int n = 0; n < NMAX; ++n) {
      if (n < p.N) {
        acc_re[n] = wave_sum(acc_re[n]);
        if constexpr (CPLX) acc_im[n] = wave_sum(acc_im[n]);
      }
    }
    T* D = p.D + (long)b * p.sD + (long)row * p.ldd;
    const T* C = p.C ? p.C + (long)b * p.sC + (long)row * p.ldc : nullptr;
#pragma unroll
    for (int n = 0; n < NMAX; ++n) {
      if (n < p.N && lane == n) {
        T v = make_val<T>(acc_re[n], acc_im[n]);
        v = s_scale(v, p.alpha);
        if (p.beta != 0.0) v = s_add(v, s_scale(C[n], p.beta));
        D[n] = v;
      }
    }
  }
}

template <typename T>
hipError_t gemm(const GemmArgs<T>& a, hipStream_t s) {
  if (a.M <= 0 || a.N <= 0 || a.batch <= 0) return hipSuccess;
  unsigned gy = (unsigned)(a.batch < 65535 ? a.batch : 65535);
  const double flops = 2.0 * a.M * a.N * (double)a.K * a.batch * flop_mult<T>();
  const double bytes = sizeof(T) * ((double)a.M * a.K * (a.sA ? a.batch : 1) + (double)a.K * a.N * (a.sB ? a.batch : 1) +
                                    (double)a.M * a.N * a.batch * (a.beta != 0.0 ? 2 : 1));
  if (a.N <= 16) {
    static const std::string nm = std::string("gemm_skinny<") + tname<T>() + ">";
    prof::Scope ps(s, nm.c_str(), flops, bytes);
    dim3 grid(ceil_div(a.M, 4), gy);
    if (a.N <= 1)
      hipLaunchKernelGGL((gemm_skinny_kernel<T, 1>), grid, dim3(256), 0, s, a);
    else if (a.N <= 4)
      hipLaunchKernelGGL((gemm_skinny_kernel<T, 4>), grid, dim3(256), 0, s, a);
    else
      hipLaunchKernelGGL((gemm_skinny_kernel<T, 16>), grid, dim3(256), 0, s, a);
  } else {
    static const std::string nm = std::string("gemm_mfma<") + tname<T>() + ">";
    prof::Scope ps(s, nm.c_str(), flops, bytes);
    int tiles_n = ceil_div(a.N, GBN);
    int tiles_m = ceil_div(a.M, GBM);
    dim3 grid((unsigned)(tiles_m * tiles_n), gy);
    hipLaunchKernelGGL((gemm_mfma_kernel<T>), grid, dim3(256), 0, s, a, tiles_n);
  }
  return hipGetLastError();
}

template hipError_t gemm<double>(const GemmArgs<double>&, hipStream_t);
template hipError_t gemm<float>(const GemmArgs<float>&, hipStream_t);
template hipError_t gemm<c128>(const GemmArgs<c128>&, hipStream_t);
template hipError_t gemm<c64>(const GemmArgs<c64>&, hipStream_t);

}  // namespace hps
