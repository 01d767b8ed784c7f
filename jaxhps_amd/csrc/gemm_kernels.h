// GEMM kernel templates for gfx950 (included by gemm.hip).
#pragma once

#include "hps_kernels.h"

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

constexpr int GBK = 16;

// Block tile (WR*TM*16) x (WC*TN*16) x 16 with WC = 4 / WR; 4 waves in a
// WR x WC arrangement, each owning TM x TN MFMA 16x16 blocks.  WR = 2 is the
// square layout; WR = 4 stacks the waves along M for N <= 16 (the
// multi-right-hand-side products).  The next K-tile is fetched from global
// memory into registers while the current one is multiplied out of LDS.
// Workgroups are handed to the 8 XCDs round-robin by linear id, and each
// XCD has its own L2.  For a batch (grid y >= 2) xcd_block gives the
// workgroup the (x, y) block of the linear id (id % 8) * (W / 8) + id / 8
// instead (W = grid size, a multiple of 8; otherwise the identity): each XCD
// then runs a contiguous run of tiles, so the tiles of one matrix (and their
// shared A row and B column panels) stay in one L2.  A single product keeps
// the round-robin order, which measured faster for it.  The map is a
// permutation of the grid.
__device__ __forceinline__ void xcd_block(int& bx, int& by) {
  const unsigned long gx = gridDim.x, w = gx * gridDim.y;
  unsigned long id = blockIdx.x + gx * blockIdx.y;
  if (gridDim.y >= 2 && w % 8 == 0) id = (id % 8) * (w / 8) + id / 8;
  bx = (int)(id % gx);
  by = (int)(id / gx);
}

template <typename T> __device__ __forceinline__ T s_conj(T x) { return x; }
template <> __device__ __forceinline__ c128 s_conj<c128>(c128 x) { return c128{x.re, -x.im}; }
template <> __device__ __forceinline__ c64 s_conj<c64>(c64 x) { return c64{x.re, -x.im}; }

// TA: op(A) = A^H with A stored K x M; the tile load then walks M fastest so
// the reads stay coalesced.
// K3M (complex only): three real products per complex one (P1 = Ar Br,
// P2 = Ai Bi, P3 = (Ar + Ai)(Br + Bi); Re = P1 - P2, Im = P3 - P1 - P2), a
// quarter fewer MFMAs for one more accumulator per block.
template <typename T, int TM, int TN, int WR, bool TA, bool K3M>
__device__ __forceinline__ void gemm_mfma_tile(const GemmArgs<T>& p, int bx, int by, int tiles_n) {
  using R = typename real_of<T>::type;
  using M = Mfma<R>;
  using acc_t = typename M::acc_t;
  constexpr bool CPLX = is_cplx<T>::value;
  constexpr int WC = 4 / WR;
  static_assert(WR * WC == 4, "four waves per block");
  constexpr int BM = WR * TM * 16, BN = WC * TN * 16;
  static_assert((BM * GBK) % 256 == 0 && (GBK * BN) % 256 == 0, "tile loads must cover the block evenly");
  constexpr int LA = BM * GBK / 256, LB = GBK * BN / 256;
  __shared__ R As_re[GBK][BM + 1];
  __shared__ R Bs_re[GBK][BN + 1];
  __shared__ R As_im[CPLX ? GBK : 1][CPLX ? BM + 1 : 1];
  __shared__ R Bs_im[CPLX ? GBK : 1][CPLX ? BN + 1 : 1];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wr = wave / WC, wc = wave % WC;
  const int tm = bx / tiles_n, tn = bx % tiles_n;
  const int m0 = tm * BM, n0 = tn * BN;
  const int li = lane & 15, lk = lane >> 4;

  for (int b = by; b < p.batch; b += gridDim.y) {
    const T* A = p.A + (long)b * p.sA;
    const T* B = p.B + (long)b * p.sB;
    static_assert(!K3M || CPLX, "three-product form is for complex types");
    acc_t acc_re[TM][TN], acc_im[TM][TN], acc_x[K3M ? TM : 1][K3M ? TN : 1];
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j) {
        acc_re[i][j] = acc_t{0, 0, 0, 0};
        acc_im[i][j] = acc_t{0, 0, 0, 0};
        if constexpr (K3M) acc_x[i][j] = acc_t{0, 0, 0, 0};
      }
    T ra[LA], rb[LB];
    auto fetch = [&](int k0) {
#pragma unroll
      for (int i = 0; i < LA; ++i) {
        int e = tid + 256 * i;
        if constexpr (TA) {
          int r = e % BM, c = e / BM;
          int gr = m0 + r, gc = k0 + c;
          ra[i] = (gr < p.M && gc < p.K) ? s_conj(A[(long)gc * p.lda + gr]) : s_zero<T>();
        } else {
          int r = e >> 4, c = e & 15;
          int gr = m0 + r, gc = k0 + c;
          ra[i] = (gr < p.M && gc < p.K) ? A[(long)gr * p.lda + gc] : s_zero<T>();
        }
      }
#pragma unroll
      for (int i = 0; i < LB; ++i) {
        int e = tid + 256 * i;
        int r = e / BN, c = e % BN;
        int gr = k0 + r, gc = n0 + c;
        rb[i] = (gr < p.K && gc < p.N) ? B[(long)gr * p.ldb + gc] : s_zero<T>();
      }
    };
    fetch(0);
    for (int k0 = 0; k0 < p.K; k0 += GBK) {
#pragma unroll
      for (int i = 0; i < LA; ++i) {
        int e = tid + 256 * i;
        int r = TA ? e % BM : e >> 4, c = TA ? e / BM : e & 15;
        As_re[c][r] = s_re(ra[i]);
        if constexpr (CPLX) As_im[c][r] = s_im(ra[i]);
      }
#pragma unroll
      for (int i = 0; i < LB; ++i) {
        int e = tid + 256 * i;
        int r = e / BN, c = e % BN;
        Bs_re[r][c] = s_re(rb[i]);
        if constexpr (CPLX) Bs_im[r][c] = s_im(rb[i]);
      }
      __syncthreads();
      if (k0 + GBK < p.K) fetch(k0 + GBK);
#pragma unroll
      for (int kk = 0; kk < GBK; kk += 4) {
        R a_re[TM], b_re[TN], a_im[TM], b_im[TN];
#pragma unroll
        for (int t = 0; t < TM; ++t) {
          a_re[t] = As_re[kk + lk][wr * (BM / WR) + t * 16 + li];
          if constexpr (CPLX) a_im[t] = As_im[kk + lk][wr * (BM / WR) + t * 16 + li];
        }
#pragma unroll
        for (int t = 0; t < TN; ++t) {
          b_re[t] = Bs_re[kk + lk][wc * (BN / WC) + t * 16 + li];
          if constexpr (CPLX) b_im[t] = Bs_im[kk + lk][wc * (BN / WC) + t * 16 + li];
        }
        if constexpr (K3M) {
          R a_s[TM], b_s[TN];
#pragma unroll
          for (int i = 0; i < TM; ++i) a_s[i] = a_re[i] + a_im[i];
#pragma unroll
          for (int j = 0; j < TN; ++j) b_s[j] = b_re[j] + b_im[j];
#pragma unroll
          for (int i = 0; i < TM; ++i)
#pragma unroll
            for (int j = 0; j < TN; ++j) {
              acc_re[i][j] = M::mma(a_re[i], b_re[j], acc_re[i][j]);
              acc_x[i][j] = M::mma(a_im[i], b_im[j], acc_x[i][j]);
              acc_im[i][j] = M::mma(a_s[i], b_s[j], acc_im[i][j]);
            }
        } else {
#pragma unroll
          for (int i = 0; i < TM; ++i)
#pragma unroll
            for (int j = 0; j < TN; ++j) {
              acc_re[i][j] = M::mma(a_re[i], b_re[j], acc_re[i][j]);
              if constexpr (CPLX) {
                acc_re[i][j] = M::mma(-a_im[i], b_im[j], acc_re[i][j]);
                acc_im[i][j] = M::mma(a_re[i], b_im[j], acc_im[i][j]);
                acc_im[i][j] = M::mma(a_im[i], b_re[j], acc_im[i][j]);
              }
            }
        }
      }
      __syncthreads();
    }
    if constexpr (K3M) {
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) {
          acc_im[i][j] = acc_im[i][j] - acc_re[i][j] - acc_x[i][j];
          acc_re[i][j] = acc_re[i][j] - acc_x[i][j];
        }
    }

    const T* C = p.C ? p.C + (long)b * p.sC : nullptr;
    T* D = p.D + (long)b * p.sD;
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          int row = m0 + wr * (BM / WR) + i * 16 + M::out_row(lk, r);
          int col = n0 + wc * (BN / WC) + j * 16 + li;
          if (row < p.M && col < p.N) {
            T v = make_val<T>(acc_re[i][j][r], CPLX ? acc_im[i][j][r] : R(0));
            v = s_scale(v, p.alpha);
            if (p.beta != 0.0) v = s_add(v, s_scale(C[(long)row * p.ldc + col], p.beta));
            D[(long)row * p.ldd + col] = v;
          }
        }
  }
}

template <typename T, int TM, int TN, int MINB = 1, int WR = 2, bool TA = false, bool K3M = false>
__global__ __launch_bounds__(256, MINB) void gemm_mfma_kernel(GemmArgs<T> p, int tiles_n) {
  int bx, by;
  xcd_block(bx, by);
  gemm_mfma_tile<T, TM, TN, WR, TA, K3M>(p, bx, by, tiles_n);
}

// Several products in one launch: group g is the product of rows
// r0[g] .. r0[g] + nr[g] of A (columns k0[g] .. k0[g] + nk[g]) with rows
// k0[g] .. of B, into rows r0[g] .. of C and D; every group shares N, the
// leading dimensions, the batch strides, alpha and beta (the block-sparse
// coupling and update products of the merges).  Workgroup x-ids
// start[g] .. start[g + 1] - 1 hold group g's tiles, so the groups fill the
// GPU together instead of one short launch after another.
constexpr int GEMM_MAX_GROUPS = 32;
struct GemmGroups {
  int n = 0;
  int start[GEMM_MAX_GROUPS + 1] = {};
  int r0[GEMM_MAX_GROUPS] = {}, nr[GEMM_MAX_GROUPS] = {}, k0[GEMM_MAX_GROUPS] = {}, nk[GEMM_MAX_GROUPS] = {};
};

template <typename T, int TM, int TN, int MINB, bool K3M>
__global__ __launch_bounds__(256, MINB) void gemm_mfma_grouped_kernel(GemmArgs<T> p, GemmGroups g, int tiles_n) {
  int bx, by;
  xcd_block(bx, by);
  int gi = 0;
  while (gi + 1 < g.n && bx >= g.start[gi + 1]) ++gi;
  bx -= g.start[gi];
  p.M = g.nr[gi];
  p.K = g.nk[gi];
  p.A += (long)g.r0[gi] * p.lda + g.k0[gi];
  p.B += (long)g.k0[gi] * p.ldb;
  if (p.C) p.C += (long)g.r0[gi] * p.ldc;
  p.D += (long)g.r0[gi] * p.ldd;
  gemm_mfma_tile<T, TM, TN, 2, false, K3M>(p, bx, by, tiles_n);
}

// Short-K updates (K <= 32: the 32-wide steps of the blocked LU and solves):
// a (32 TM) x (32 TN) tile whose whole A panel (BM x K), B panel (K x BN) and
// C tile are requested at once, one LDS pass, one barrier, then the MFMAs and
// the epilogue from the prefetched C.  Four waves in a 2 x 2 arrangement, each
// owning TM x TN 16 x 16 blocks.  TRI = 1 / 2 reads A as the unit-lower /
// upper part of a packed diagonal-block inverse (the triangular solves' block
// step, M <= 32: TM = 1); every B element of a tile is read before any D
// element is written, so D may alias B.
template <typename T, int TRI = 0, int TM = 2, int TN = 1>
__global__ __launch_bounds__(256, 2) void gemm_k32_kernel(GemmArgs<T> p, int tiles_n) {
  using R = typename real_of<T>::type;
  using M = Mfma<R>;
  using acc_t = typename M::acc_t;
  constexpr bool CPLX = is_cplx<T>::value;
  constexpr int BM = 32 * TM, BN = 32 * TN, KB = 32;
  constexpr int LA = BM * KB / 256, LB = KB * BN / 256;
  __shared__ R As_re[KB][BM + 1];
  __shared__ R Bs_re[KB][BN + 1];
  __shared__ R As_im[CPLX ? KB : 1][CPLX ? BM + 1 : 1];
  __shared__ R Bs_im[CPLX ? KB : 1][CPLX ? BN + 1 : 1];
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wr = wave >> 1, wc = wave & 1;
  int bx, by;
  xcd_block(bx, by);
  const int tm = bx / tiles_n, tn = bx % tiles_n;
  const int m0 = tm * BM, n0 = tn * BN;
  const int li = lane & 15, lk = lane >> 4;
  for (int b = by; b < p.batch; b += gridDim.y) {
    const T* A = p.A + (long)b * p.sA;
    const T* B = p.B + (long)b * p.sB;
    const T* C = p.C ? p.C + (long)b * p.sC : nullptr;
    T* D = p.D + (long)b * p.sD;
    T ra[LA], rb[LB], rc[TM][TN][4];
#pragma unroll
    for (int i = 0; i < LA; ++i) {
      const int e = tid + 256 * i, r = e >> 5, c = e & 31;
      const int gr = m0 + r;
      T v = s_zero<T>();
      if (gr < p.M && c < p.K) {
        if (TRI == 1)
          v = c < gr ? A[(long)gr * p.lda + c] : (c == gr ? s_one<T>() : s_zero<T>());
        else if (TRI == 2)
          v = c >= gr ? A[(long)gr * p.lda + c] : s_zero<T>();
        else
          v = A[(long)gr * p.lda + c];
      }
      ra[i] = v;
    }
#pragma unroll
    for (int i = 0; i < LB; ++i) {
      const int e = tid + 256 * i, r = e / BN, c = e % BN;
      const int gc = n0 + c;
      rb[i] = (r < p.K && gc < p.N) ? B[(long)r * p.ldb + gc] : s_zero<T>();
    }
    const bool use_c = p.beta != 0.0;
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = m0 + wr * (BM / 2) + i * 16 + M::out_row(lk, r);
          const int col = n0 + wc * (BN / 2) + j * 16 + li;
          rc[i][j][r] = (use_c && row < p.M && col < p.N) ? C[(long)row * p.ldc + col] : s_zero<T>();
        }
#pragma unroll
    for (int i = 0; i < LA; ++i) {
      const int e = tid + 256 * i, r = e >> 5, c = e & 31;
      As_re[c][r] = s_re(ra[i]);
      if constexpr (CPLX) As_im[c][r] = s_im(ra[i]);
    }
#pragma unroll
    for (int i = 0; i < LB; ++i) {
      const int e = tid + 256 * i, r = e / BN, c = e % BN;
      Bs_re[r][c] = s_re(rb[i]);
      if constexpr (CPLX) Bs_im[r][c] = s_im(rb[i]);
    }
    __syncthreads();
    acc_t acc_re[TM][TN], acc_im[TM][TN];
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j) {
        acc_re[i][j] = acc_t{0, 0, 0, 0};
        acc_im[i][j] = acc_t{0, 0, 0, 0};
      }
#pragma unroll
    for (int kk = 0; kk < KB; kk += 4) {
      R a_re[TM], a_im[TM], b_re[TN], b_im[TN];
#pragma unroll
      for (int i = 0; i < TM; ++i) {
        a_re[i] = As_re[kk + lk][wr * (BM / 2) + i * 16 + li];
        a_im[i] = CPLX ? As_im[kk + lk][wr * (BM / 2) + i * 16 + li] : R(0);
      }
#pragma unroll
      for (int j = 0; j < TN; ++j) {
        b_re[j] = Bs_re[kk + lk][wc * (BN / 2) + j * 16 + li];
        b_im[j] = CPLX ? Bs_im[kk + lk][wc * (BN / 2) + j * 16 + li] : R(0);
      }
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) {
          acc_re[i][j] = M::mma(a_re[i], b_re[j], acc_re[i][j]);
          if constexpr (CPLX) {
            acc_re[i][j] = M::mma(-a_im[i], b_im[j], acc_re[i][j]);
            acc_im[i][j] = M::mma(a_re[i], b_im[j], acc_im[i][j]);
            acc_im[i][j] = M::mma(a_im[i], b_re[j], acc_im[i][j]);
          }
        }
    }
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
      for (int j = 0; j < TN; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = m0 + wr * (BM / 2) + i * 16 + M::out_row(lk, r);
          const int col = n0 + wc * (BN / 2) + j * 16 + li;
          if (row < p.M && col < p.N) {
            T v = make_val<T>(acc_re[i][j][r], CPLX ? acc_im[i][j][r] : R(0));
            v = s_scale(v, p.alpha);
            if (use_c) v = s_add(v, s_scale(rc[i][j][r], p.beta));
            D[(long)row * p.ldd + col] = v;
          }
        }
    // the next batch element rewrites the tiles
    __syncthreads();
  }
}

template <typename R>
__device__ __forceinline__ R wave_sum(R v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// N <= 2 with op(A) = A^H (A stored K x M): thread m accumulates column m of
// A against B; consecutive threads read consecutive elements of A's rows.
template <typename T, int NMAX>
__global__ __launch_bounds__(256) void gemv_conjt_kernel(GemmArgs<T> p) {
  const int m = blockIdx.x * 256 + threadIdx.x;
  if (m >= p.M) return;
  for (int b = blockIdx.y; b < p.batch; b += gridDim.y) {
    const T* A = p.A + (long)b * p.sA + m;
    const T* B = p.B + (long)b * p.sB;
    T acc[NMAX];
#pragma unroll
    for (int n = 0; n < NMAX; ++n) acc[n] = s_zero<T>();
    for (int k = 0; k < p.K; ++k) {
      const T a = s_conj(A[(long)k * p.lda]);
#pragma unroll
      for (int n = 0; n < NMAX; ++n)
        if (n < p.N) acc[n] = s_add(acc[n], s_mul(a, B[(long)k * p.ldb + n]));
    }
    T* D = p.D + (long)b * p.sD + (long)m * p.ldd;
    const T* C = p.C ? p.C + (long)b * p.sC + (long)m * p.ldc : nullptr;
#pragma unroll
    for (int n = 0; n < NMAX; ++n) {
      if (n < p.N) {
        T v = s_scale(acc[n], p.alpha);
        if (p.beta != 0.0) v = s_add(v, s_scale(C[n], p.beta));
        D[n] = v;
      }
    }
  }
}

// N <= 2: LPR lanes per output row (64 = one wave per row for long rows;
// 16 or 4 for the short rows of the leaf apply and the lower tree levels, so
// a wave serves 4 or 16 rows and the cross-lane sum is 4 or 2 steps).  Each
// row's lanes read it contiguously.  256 / LPR rows per workgroup.
template <typename T, int NMAX, int LPR = 64>
__global__ __launch_bounds__(256) void gemm_skinny_kernel(GemmArgs<T> p) {
  using R = typename real_of<T>::type;
  constexpr bool CPLX = is_cplx<T>::value;
  const int sub = threadIdx.x % LPR;
  const int row = blockIdx.x * (256 / LPR) + threadIdx.x / LPR;
  const bool live = row < p.M;  // every lane takes part in the cross-lane sums
  for (int b = blockIdx.y; b < p.batch; b += gridDim.y) {
    const T* A = p.A + (long)b * p.sA + (long)(live ? row : 0) * p.lda;
    const T* B = p.B + (long)b * p.sB;
    R acc_re[NMAX], acc_im[NMAX];
#pragma unroll
    for (int n = 0; n < NMAX; ++n) {
      acc_re[n] = R(0);
      acc_im[n] = R(0);
    }
    auto fma_into = [&](const T& a, const T& bv, int n) {
      if constexpr (CPLX) {
        acc_re[n] += s_re(a) * s_re(bv) - s_im(a) * s_im(bv);
        acc_im[n] += s_re(a) * s_im(bv) + s_im(a) * s_re(bv);
      } else {
        acc_re[n] += s_re(a) * s_re(bv);
      }
    };
    // four k steps per iteration: their loads are all issued before the
    // products, so each lane has four rows' worth of A and B in flight
    int k = sub;
    for (; k + 3 * LPR < p.K && live; k += 4 * LPR) {
      T a[4], bv[4][NMAX];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        a[u] = A[k + u * LPR];
        const T* Bk = B + (long)(k + u * LPR) * p.ldb;
#pragma unroll
        for (int n = 0; n < NMAX; ++n) bv[u][n] = n < p.N ? Bk[n] : s_zero<T>();
      }
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int n = 0; n < NMAX; ++n)
          if (n < p.N) fma_into(a[u], bv[u][n], n);
    }
    for (; k < p.K && live; k += LPR) {
      T a = A[k];
      const T* Bk = B + (long)k * p.ldb;
#pragma unroll
      for (int n = 0; n < NMAX; ++n)
        if (n < p.N) fma_into(a, Bk[n], n);
    }
#pragma unroll
    for (int n = 0; n < NMAX; ++n) {
      if (n < p.N) {
#pragma unroll
        for (int off = LPR / 2; off > 0; off >>= 1) {
          acc_re[n] += __shfl_xor(acc_re[n], off, 64);
          if constexpr (CPLX) acc_im[n] += __shfl_xor(acc_im[n], off, 64);
        }
      }
    }
    if (!live) continue;
    T* D = p.D + (long)b * p.sD + (long)row * p.ldd;
    const T* C = p.C ? p.C + (long)b * p.sC + (long)row * p.ldc : nullptr;
#pragma unroll
    for (int n = 0; n < NMAX; ++n) {
      if (n < p.N && sub == n) {
        T v = make_val<T>(acc_re[n], acc_im[n]);
        v = s_scale(v, p.alpha);
        if (p.beta != 0.0) v = s_add(v, s_scale(C[n], p.beta));
        D[n] = v;
      }
    }
  }
}


}  // namespace hps
