// Exterior-scattering coupling of the top-level operator
// (examples/wave_scattering_utils.py:31-244): incident plane waves and their
// outward normal derivatives on the boundary, ItI -> DtN conversion, the
// boundary-integral system A = I/2 - D + S T, b = S (du_in/dn - T u_in) and
// the incoming impedance data of the scattered field.
//
// Everything is one dense complex128 problem on the n root-boundary points:
// the O(n^3) work runs in the library's LU (getrf_ws / getrs_perm) and GEMM
// kernels; the O(n^2) and O(n nsrc) glue is the two small kernels below.
#include <string>

#include "hps_capi_internal.h"

namespace hps {

// u_in[i, s] = exp(i k <x_i, d_s>), d_s = (cos th_s, sin th_s).  With dn,
// dn[i, s] = d u_in / dn on the root boundary, whose points come in four sides
// S, E, N, W of n_per_side points (the last side takes any remainder) with
// outward normals (0,-1), (1,0), (0,1), (-1,0).
__global__ __launch_bounds__(256) void plane_wave_kernel(long npts, int nsrc, const double* pts, const double* dirs,
                                                         double k, long n_per_side, c128* uin, c128* dn) {
  const long total = npts * nsrc;
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < total; t += (long)gridDim.x * blockDim.x) {
    const long i = t / nsrc;
    const int s = (int)(t - i * nsrc);
    double sx, cx;
    sincos(dirs[s], &sx, &cx);
    const double ph = k * (pts[2 * i] * cx + pts[2 * i + 1] * sx);
    double sp, cp;
    sincos(ph, &sp, &cp);
    uin[t] = c128{cp, sp};
    if (dn) {
      long side = i / n_per_side;
      if (side > 3) side = 3;
      // outward normal component of the wave vector: -y, +x, +y, -x
      const double kn = side == 0 ? -k * sx : side == 1 ? k * cx : side == 2 ? k * sx : -k * cx;
      // i kn u_in
      dn[t] = c128{-kn * sp, kn * cp};
    }
  }
}

// out[r, c] = a X[r, c] + b Y[r, c] + (r == c ? d : 0); X or Y may be null
// (their term is dropped); out may alias X or Y.
__global__ __launch_bounds__(256) void combine_kernel(int rows, int cols, c128 a, const c128* X, long ldx, c128 b,
                                                      const c128* Y, long ldy, c128 d, c128* out, long ldo) {
  const long total = (long)rows * cols;
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < total; t += (long)gridDim.x * blockDim.x) {
    const int r = (int)(t / cols), c = (int)(t - (long)r * cols);
    c128 v = r == c ? d : s_zero<c128>();
    if (X) v = s_fma(a, X[r * ldx + c], v);
    if (Y) v = s_fma(b, Y[r * ldy + c], v);
    out[r * ldo + c] = v;
  }
}

static hipError_t combine(int rows, int cols, c128 a, const c128* X, long ldx, c128 b, const c128* Y, long ldy, c128 d,
                          c128* out, long ldo, hipStream_t s) {
  if ((long)rows * cols == 0) return hipSuccess;
  hipLaunchKernelGGL(combine_kernel, grid_for((long)rows * cols, 1), dim3(256), 0, s, rows, cols, a, X, ldx, b, Y, ldy,
                     d, out, ldo);
  return hipGetLastError();
}

static hipError_t plane_wave(long npts, int nsrc, const double* pts, const double* dirs, double k, long n_per_side,
                             c128* uin, c128* dn, hipStream_t s) {
  if (npts * nsrc == 0) return hipSuccess;
  hipLaunchKernelGGL(plane_wave_kernel, grid_for(npts * nsrc, 1), dim3(256), 0, s, npts, nsrc, pts, dirs, k,
                     n_per_side, uin, dn);
  return hipGetLastError();
}

// dense n x n solve of one matrix: A is factored in place, X = A^{-1} R
struct DenseLU {
  int* perm;
  c128* dinv;
  int* scratch;
};

static size_t dense_lu_bytes(int n) {
  return align_up((size_t)n * sizeof(int)) + align_up(dinv_elems(n) * sizeof(c128)) +
         align_up(getrf_scratch_bytes(n, 1));
}

static DenseLU carve_lu(char*& p, int n) {
  DenseLU f;
  f.perm = (int*)p;
  p += align_up((size_t)n * sizeof(int));
  f.dinv = (c128*)p;
  p += align_up(dinv_elems(n) * sizeof(c128));
  f.scratch = (int*)p;
  p += align_up(getrf_scratch_bytes(n, 1));
  return f;
}

static size_t mat_bytes(long rows, long cols) { return align_up((size_t)rows * cols * sizeof(c128)); }

static c128* carve(char*& p, long rows, long cols) {
  c128* out = (c128*)p;
  p += mat_bytes(rows, cols);
  return out;
}

static hipError_t gemm_c128(int M, int N, int K, double alpha, const c128* A, const c128* B, double beta,
                            const c128* C, c128* D, hipStream_t s) {
  GemmArgs<c128> g;
  g.M = M;
  g.N = N;
  g.K = K;
  g.alpha = alpha;
  g.beta = beta;
  g.A = A;
  g.lda = K;
  g.B = B;
  g.ldb = N;
  g.C = C;
  g.ldc = N;
  g.D = D;
  g.ldd = N;
  return gemm<c128>(g, s);
}

// A = I/2 - D + S T and b = S (dn - T u_in), with u_in and dn left in uin / dn
static hipError_t scattering_system(int n, int nsrc, const c128* S, const c128* D, const c128* T, const double* pts,
                                    const double* dirs, double k, c128* A, c128* b, c128* uin, c128* dn, c128* r,
                                    hipStream_t s) {
  hipError_t e;
  if ((e = plane_wave(n, nsrc, pts, dirs, k, n / 4, uin, dn, s)) != hipSuccess) return e;
  if ((e = gemm_c128(n, n, n, 1.0, S, T, -1.0, D, A, s)) != hipSuccess) return e;
  const c128 one = s_one<c128>(), zero = s_zero<c128>();
  if ((e = combine(n, n, one, A, n, zero, nullptr, 0, c128{0.5, 0.0}, A, n, s)) != hipSuccess) return e;
  if ((e = gemm_c128(n, nsrc, n, -1.0, T, uin, 1.0, dn, r, s)) != hipSuccess) return e;
  return gemm_c128(n, nsrc, n, 1.0, S, r, 0.0, nullptr, b, s);
}

}  // namespace hps

using namespace hps;

static int require_c128(hps_handle_t h, int dtype) {
  if (dtype != HPS_C128) {
    h->err = "the scattering coupling is complex128 (HPS_C128) only";
    return HPS_ERR_ARG;
  }
  return HPS_OK;
}

int hps_plane_wave(hps_handle_t h, long npts, int nsrc, const double* pts, const double* dirs, double k,
                   long n_per_side, void* uin, void* uin_dn) {
  if (!h) return HPS_ERR_ARG;
  HPS_ARG(h, npts >= 0 && nsrc >= 0, "bad size");
  if (npts * nsrc == 0) return HPS_OK;
  HPS_ARG(h, pts && dirs && uin, "null pointer");
  HPS_ARG(h, !uin_dn || n_per_side > 0, "normal derivatives need n_per_side > 0");
  hipStream_t s = h->stream;
  HPS_PROF("plane_wave", c128, 0.0, (double)npts * nsrc * (uin_dn ? 32.0 : 16.0) + npts * 16.0);
  HPS_CHECK_HIP(h, plane_wave(npts, nsrc, pts, dirs, k, n_per_side, (c128*)uin, (c128*)uin_dn, s));
  return HPS_OK;
}

size_t hps_iti_to_dtn_workspace(int n) { return 2 * mat_bytes(n, n) + dense_lu_bytes(n); }

int hps_iti_to_dtn(hps_handle_t h, int dtype, int n, const void* R, long ldr, double eta, void* T, long ldt,
                   int* info, void* ws, size_t ws_bytes) {
  if (!h) return HPS_ERR_ARG;
  if (int rc = require_c128(h, dtype)) return rc;
  HPS_ARG(h, n >= 0 && n <= 12288 && ldr >= n && ldt >= n, "bad size (0 <= n <= 12288, ldr, ldt >= n)");
  if (n == 0) return HPS_OK;
  HPS_ARG(h, R && T && info && ws, "null pointer");
  if (ws_bytes < hps_iti_to_dtn_workspace(n)) {
    h->err = "iti_to_dtn workspace too small";
    return HPS_ERR_WORKSPACE;
  }
  hipStream_t s = h->stream;
  char* p = (char*)ws;
  c128* W = carve(p, n, n);
  c128* B = carve(p, n, n);
  DenseLU f = carve_lu(p, n);
  const c128* Rc = (const c128*)R;
  const c128 one = s_one<c128>(), zero = s_zero<c128>(), mie = c128{0.0, -eta};
  // W = R - I and B = -i eta (R + I): T = W^{-1} B.  R is read before T is
  // written, so T may alias R.
  HPS_CHECK_HIP(h, combine(n, n, one, Rc, ldr, zero, nullptr, 0, c128{-1.0, 0.0}, W, n, s));
  HPS_CHECK_HIP(h, combine(n, n, mie, Rc, ldr, zero, nullptr, 0, mie, B, n, s));
  HPS_CHECK_HIP(h, getrf_ws<c128>(n, W, n, 0, f.perm, f.dinv, f.scratch, info, 1, s));
  HPS_CHECK_HIP(h, getrs_perm<c128>(n, n, W, n, 0, f.perm, f.dinv, B, n, 0, (c128*)T, ldt, 0, 1, s));
  return HPS_OK;
}

size_t hps_scattering_workspace(int n, int nsrc) { return mat_bytes(n, n) + 5 * mat_bytes(n, nsrc) + dense_lu_bytes(n); }

int hps_scattering_system(hps_handle_t h, int dtype, int n, int nsrc, const void* S, const void* D, const void* T,
                          const double* bdry_pts, const double* source_dirs, double k, void* A, void* b, void* ws,
                          size_t ws_bytes) {
  if (!h) return HPS_ERR_ARG;
  if (int rc = require_c128(h, dtype)) return rc;
  HPS_ARG(h, n >= 4 && nsrc >= 1, "bad size (n >= 4 boundary points, nsrc >= 1)");
  HPS_ARG(h, S && D && T && bdry_pts && source_dirs && A && b && ws, "null pointer");
  HPS_ARG(h, A != S && A != D && A != T, "A must not alias S, D or T");
  if (ws_bytes < hps_scattering_workspace(n, nsrc)) {
    h->err = "scattering workspace too small";
    return HPS_ERR_WORKSPACE;
  }
  hipStream_t s = h->stream;
  char* p = (char*)ws;
  p += mat_bytes(n, n);
  c128* uin = carve(p, n, nsrc);
  c128* dn = carve(p, n, nsrc);
  c128* r = carve(p, n, nsrc);
  HPS_CHECK_HIP(h, scattering_system(n, nsrc, (const c128*)S, (const c128*)D, (const c128*)T, bdry_pts, source_dirs,
                                     k, (c128*)A, (c128*)b, uin, dn, r, s));
  return HPS_OK;
}

int hps_scattering_impedance(hps_handle_t h, int dtype, int n, int nsrc, const void* S, const void* D, const void* T,
                             const double* bdry_pts, const double* source_dirs, double k, double eta, void* imp,
                             int* info, void* ws, size_t ws_bytes) {
  if (!h) return HPS_ERR_ARG;
  if (int rc = require_c128(h, dtype)) return rc;
  HPS_ARG(h, n >= 4 && n <= 12288 && nsrc >= 1, "bad size (4 <= n <= 12288 boundary points, nsrc >= 1)");
  HPS_ARG(h, S && D && T && bdry_pts && source_dirs && imp && info && ws, "null pointer");
  if (ws_bytes < hps_scattering_workspace(n, nsrc)) {
    h->err = "scattering workspace too small";
    return HPS_ERR_WORKSPACE;
  }
  hipStream_t s = h->stream;
  char* p = (char*)ws;
  c128* A = carve(p, n, n);
  c128* uin = carve(p, n, nsrc);
  c128* dn = carve(p, n, nsrc);
  c128* r = carve(p, n, nsrc);
  c128* b = carve(p, n, nsrc);
  c128* us = carve(p, n, nsrc);
  DenseLU f = carve_lu(p, n);
  const c128* Tc = (const c128*)T;
  const c128 one = s_one<c128>(), zero = s_zero<c128>();
  HPS_CHECK_HIP(h, scattering_system(n, nsrc, (const c128*)S, (const c128*)D, Tc, bdry_pts, source_dirs, k, A, b, uin,
                                     dn, r, s));
  // u_scat = A^{-1} b on the boundary
  HPS_CHECK_HIP(h, getrf_ws<c128>(n, A, n, 0, f.perm, f.dinv, f.scratch, info, 1, s));
  HPS_CHECK_HIP(h, getrs_perm<c128>(n, nsrc, A, n, 0, f.perm, f.dinv, b, nsrc, 0, us, nsrc, 0, 1, s));
  // d u_scat / dn = T (u_scat + u_in) - d u_in / dn, then imp = du_scat/dn + i eta u_scat
  HPS_CHECK_HIP(h, combine(n, nsrc, one, us, nsrc, one, uin, nsrc, zero, r, nsrc, s));
  HPS_CHECK_HIP(h, gemm_c128(n, nsrc, n, 1.0, Tc, r, -1.0, dn, b, s));
  HPS_CHECK_HIP(h, combine(n, nsrc, one, b, nsrc, c128{0.0, eta}, us, nsrc, zero, (c128*)imp, nsrc, s));
  return HPS_OK;
}
