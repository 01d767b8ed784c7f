// The learned leaf-ItI surrogate: kernels for training and inference of a
// small MLP (q samples -> rank-r complex correction U V^H of the q = 0 leaf
// ItI map).  No counterpart in the reference.
//
// * mlp_gemm_kernel<TA, TB>: C = op(A) op(B) (+ bias, tanh) with A, B held in
//   FP32 memory, rounded to BF16 when a tile is staged into LDS and
//   multiplied with v_mfma_f32_16x16x32_bf16 (FP32 accumulation).  64x64x32
//   block tile, 4 waves in 2x2, each wave 2x2 MFMA 16x16 blocks.  The three
//   layouts cover forward (X W^T), input gradient (dZ W) and weight gradient
//   (dZ^T X).
// * elementwise: tanh backward, bias gradient (column sums), Adam.
// * low-rank loss: E = U V^H - dT per sample, loss |E|_F^2 and its gradients
//   dU = 2 E V, dV = 2 E^H U (torch's convention for real losses of complex
//   values), and T = T0 + U V^H for inference.
#include "hps_capi_internal.h"

namespace hps {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));

constexpr int MB_M = 64, MB_N = 64, MB_K = 32;

// C[m][n] = sum_k opA[m][k] opB[k][n]; opA = TA ? A^T : A (A stored K x M or M x K),
// opB = TB ? B^T : B.  Epilogue: + bias[n] and tanh when act.
template <bool TA, bool TB>
__global__ __launch_bounds__(256) void mlp_gemm_kernel(int M, int N, int K, const float* __restrict__ A, int lda,
                                                       const float* __restrict__ B, int ldb, float* __restrict__ C,
                                                       int ldc, const float* __restrict__ bias, int act, float beta) {
  __shared__ __bf16 As[MB_M][MB_K + 8];  // [m][k]
  __shared__ __bf16 Bs[MB_N][MB_K + 8];  // [n][k]
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = wave >> 1, wc = wave & 1;
  const int tiles_n = (N + MB_N - 1) / MB_N;
  const int m0 = (blockIdx.x / tiles_n) * MB_M, n0 = (blockIdx.x % tiles_n) * MB_N;
  const int li = lane & 15, lg = lane >> 4;
  f32x4_t acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = f32x4_t{0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < K; k0 += MB_K) {
    // stage 64 x 32 of opA and 64 x 32 of opB^T (each 2048 values, 8 per thread)
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int idx = tid + 256 * e;
      int r, c;
      if (TA) {  // walk m fastest: A stored K x M
        r = idx % MB_M;
        c = idx / MB_M;
      } else {
        r = idx / MB_K;
        c = idx % MB_K;
      }
      const int gm = m0 + r, gk = k0 + c;
      float v = 0.f;
      if (gm < M && gk < K) v = TA ? A[(long)gk * lda + gm] : A[(long)gm * lda + gk];
      As[r][c] = (__bf16)v;
    }
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int idx = tid + 256 * e;
      int r, c;
      if (TB) {  // B stored N x K: walk k fastest
        r = idx / MB_K;
        c = idx % MB_K;
      } else {  // B stored K x N: walk n fastest
        r = idx % MB_N;
        c = idx / MB_N;
      }
      const int gn = n0 + r, gk = k0 + c;
      float v = 0.f;
      if (gn < N && gk < K) v = TB ? B[(long)gn * ldb + gk] : B[(long)gk * ldb + gn];
      Bs[r][c] = (__bf16)v;
    }
    __syncthreads();
    bf16x8_t a[2], b[2];
#pragma unroll
    for (int t = 0; t < 2; ++t) {
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        a[t][q] = As[wr * 32 + t * 16 + li][lg * 8 + q];
        b[t][q] = Bs[wc * 32 + t * 16 + li][lg * 8 + q];
      }
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = m0 + wr * 32 + i * 16 + 4 * lg + r;
        const int col = n0 + wc * 32 + j * 16 + li;
        if (row < M && col < N) {
          float v = acc[i][j][r];
          if (beta != 0.f) v += beta * C[(long)row * ldc + col];
          if (bias) v += bias[col];
          if (act) v = tanhf(v);
          C[(long)row * ldc + col] = v;
        }
      }
}

// dZ = dH * (1 - H^2) (tanh), in place on dH
__global__ void tanh_backward_kernel(long n, const float* __restrict__ H, float* __restrict__ dH) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const float h = H[i];
    dH[i] *= 1.f - h * h;
  }
}

// db[n] = sum_m dZ[m][n]
__global__ void bias_grad_kernel(int M, int N, const float* __restrict__ dZ, float* __restrict__ db) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  float s = 0.f;
  for (int m = 0; m < M; ++m) s += dZ[(long)m * N + n];
  db[n] = s;
}

__global__ void adam_kernel(long n, float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m1,
                            float* __restrict__ m2, float lr, float b1, float b2, float eps, float c1, float c2) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const float gi = g[i];
    const float a = b1 * m1[i] + (1.f - b1) * gi;
    const float b = b2 * m2[i] + (1.f - b2) * gi * gi;
    m1[i] = a;
    m2[i] = b;
    w[i] -= lr * (a / c1) / (sqrtf(b / c2) + eps);
  }
}

// Output row layout per sample (4 * nq * r values): U re, U im, V re, V im,
// each nq x r row-major.  One workgroup per sample; E = U V^H - dT in LDS.
__global__ __launch_bounds__(256) void lowrank_loss_kernel(int nq, int r, const float* __restrict__ Y,
                                                           const float* __restrict__ dT, float scale,
                                                           float* __restrict__ dY, float* __restrict__ loss) {
  extern __shared__ float sm[];
  const int s = blockIdx.x, tid = threadIdx.x;
  const int ur = nq * r;
  const float* y = Y + (long)s * 4 * ur;
  float* Ere = sm;
  float* Eim = sm + nq * nq;
  float* red = sm + 2 * nq * nq;
  const float* t = dT + (long)s * 2 * nq * nq;  // (nq, nq) complex interleaved
  float part = 0.f;
  for (int idx = tid; idx < nq * nq; idx += 256) {
    const int i = idx / nq, j = idx - i * nq;
    float re = -t[2 * idx], im = -t[2 * idx + 1];
    for (int q = 0; q < r; ++q) {
      const float ur_ = y[i * r + q], ui = y[ur + i * r + q];
      const float vr = y[2 * ur + j * r + q], vi = y[3 * ur + j * r + q];
      // U V^H: u * conj(v)
      re += ur_ * vr + ui * vi;
      im += ui * vr - ur_ * vi;
    }
    Ere[idx] = re;
    Eim[idx] = im;
    part += re * re + im * im;
  }
  for (int o = 32; o > 0; o >>= 1) part += __shfl_down(part, o, 64);
  if ((tid & 63) == 0) red[tid >> 6] = part;
  __syncthreads();
  if (tid == 0) loss[s] = scale * (red[0] + red[1] + red[2] + red[3]);
  float* dy = dY + (long)s * 4 * ur;
  for (int idx = tid; idx < ur; idx += 256) {
    const int i = idx / r, q = idx - i * r;
    // dU[i][q] = 2 sum_j E[i][j] v[j][q];  dV[i][q] = 2 sum_j conj(E[j][i]) u[j][q]
    float gur = 0.f, gui = 0.f, gvr = 0.f, gvi = 0.f;
    for (int j = 0; j < nq; ++j) {
      const float er = Ere[i * nq + j], ei = Eim[i * nq + j];
      const float vr = y[2 * ur + j * r + q], vi = y[3 * ur + j * r + q];
      gur += er * vr - ei * vi;
      gui += er * vi + ei * vr;
      const float fr = Ere[j * nq + i], fi = -Eim[j * nq + i];
      const float uj_r = y[j * r + q], uj_i = y[ur + j * r + q];
      gvr += fr * uj_r - fi * uj_i;
      gvi += fr * uj_i + fi * uj_r;
    }
    dy[idx] = 2.f * scale * gur;
    dy[ur + idx] = 2.f * scale * gui;
    dy[2 * ur + idx] = 2.f * scale * gvr;
    dy[3 * ur + idx] = 2.f * scale * gvi;
  }
}

// T[s] = T0 + out_scale * U V^H  (complex128 output, interleaved)
__global__ __launch_bounds__(256) void lowrank_apply_kernel(int nq, int r, const float* __restrict__ Y,
                                                            const double* __restrict__ T0, double out_scale,
                                                            double* __restrict__ T) {
  const int s = blockIdx.x;
  const int ur = nq * r;
  const float* y = Y + (long)s * 4 * ur;
  double* o = T + (long)s * 2 * nq * nq;
  for (int idx = threadIdx.x; idx < nq * nq; idx += blockDim.x) {
    const int i = idx / nq, j = idx - i * nq;
    double re = 0.0, im = 0.0;
    for (int q = 0; q < r; ++q) {
      const double ur_ = y[i * r + q], ui = y[ur + i * r + q];
      const double vr = y[2 * ur + j * r + q], vi = y[3 * ur + j * r + q];
      re += ur_ * vr + ui * vi;
      im += ui * vr - ur_ * vi;
    }
    o[2 * idx] = T0[2 * idx] + out_scale * re;
    o[2 * idx + 1] = T0[2 * idx + 1] + out_scale * im;
  }
}

}  // namespace hps

// ===========================================================================
// C ABI
// ===========================================================================
int hps_mlp_gemm_bf16(hps_handle_t h, int transA, int transB, int M, int N, int K, const float* A, int lda,
                      const float* B, int ldb, float beta, float* C, int ldc, const float* bias, int act) {
  if (!h) return HPS_ERR_ARG;
  HPS_ARG(h, M > 0 && N > 0 && K > 0, "bad GEMM sizes");
  HPS_ARG(h, A && B && C, "null pointer");
  hipStream_t s = h->stream;
  const double flops = 2.0 * M * N * (double)K;
  prof::Scope ps(s, "mlp_gemm_bf16", flops, 4.0 * ((double)M * K + (double)K * N + (double)M * N));
  const unsigned grid = ceil_div(M, MB_M) * ceil_div(N, MB_N);
#define MLP_LAUNCH(TA, TB)                                                                                       \
  hipLaunchKernelGGL((mlp_gemm_kernel<TA, TB>), dim3(grid), dim3(256), 0, s, M, N, K, A, lda, B, ldb, C, ldc, bias, \
                     act, beta)
  if (!transA && !transB)
    MLP_LAUNCH(false, false);
  else if (!transA && transB)
    MLP_LAUNCH(false, true);
  else if (transA && !transB)
    MLP_LAUNCH(true, false);
  else
    MLP_LAUNCH(true, true);
#undef MLP_LAUNCH
  HPS_CHECK_HIP(h, hipGetLastError());
  return HPS_OK;
}

int hps_mlp_tanh_backward(hps_handle_t h, long n, const float* H, float* dH) {
  if (!h) return HPS_ERR_ARG;
  HPS_ARG(h, n >= 0 && (n == 0 || (H && dH)), "bad arguments");
  if (n == 0) return HPS_OK;
  hipStream_t s = h->stream;
  long blocks = (n + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(tanh_backward_kernel, dim3((unsigned)blocks), dim3(256), 0, s, n, H, dH);
  HPS_CHECK_HIP(h, hipGetLastError());
  return HPS_OK;
}

int hps_mlp_bias_grad(hps_handle_t h, int M, int N, const float* dZ, float* db) {
  if (!h) return HPS_ERR_ARG;
  HPS_ARG(h, M > 0 && N > 0 && dZ && db, "bad arguments");
  hipStream_t s = h->stream;
  hipLaunchKernelGGL(bias_grad_kernel, dim3(ceil_div(N, 256)), dim3(256), 0, s, M, N, dZ, db);
  HPS_CHECK_HIP(h, hipGetLastError());
  return HPS_OK;
}

int hps_mlp_adam(hps_handle_t h, long n, float* w, const float* g, float* m1, float* m2, float lr, float b1, float b2,
                 float eps, int step) {
  if (!h) return HPS_ERR_ARG;
  HPS_ARG(h, n > 0 && w && g && m1 && m2 && step >= 1, "bad arguments");
  hipStream_t s = h->stream;
  const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
  long blocks = (n + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(adam_kernel, dim3((unsigned)blocks), dim3(256), 0, s, n, w, g, m1, m2, lr, b1, b2, eps, c1, c2);
  HPS_CHECK_HIP(h, hipGetLastError());
  return HPS_OK;
}

int hps_mlp_lowrank_loss(hps_handle_t h, int nsamp, int nq, int r, const float* Y, const float* dT, float scale,
                         float* dY, float* loss) {
  if (!h) return HPS_ERR_ARG;
  HPS_ARG(h, nsamp > 0 && nq > 0 && nq <= 64 && r > 0 && r <= 16, "bad sizes (nq <= 64, r <= 16)");
  HPS_ARG(h, Y && dT && dY && loss, "null pointer");
  hipStream_t s = h->stream;
  const size_t shm = (2 * (size_t)nq * nq + 8) * sizeof(float);
  hipLaunchKernelGGL(lowrank_loss_kernel, dim3(nsamp), dim3(256), shm, s, nq, r, Y, dT, scale, dY, loss);
  HPS_CHECK_HIP(h, hipGetLastError());
  return HPS_OK;
}

int hps_mlp_lowrank_apply(hps_handle_t h, int nsamp, int nq, int r, const float* Y, const double* T0,
                          double out_scale, double* T) {
  if (!h) return HPS_ERR_ARG;
  HPS_ARG(h, nsamp > 0 && nq > 0 && r > 0 && Y && T0 && T, "bad arguments");
  hipStream_t s = h->stream;
  hipLaunchKernelGGL(lowrank_apply_kernel, dim3(nsamp), dim3(256), 0, s, nq, r, Y, T0, out_scale, T);
  HPS_CHECK_HIP(h, hipGetLastError());
  return HPS_OK;
}
