// Interpolation from the HPS leaf grids to arbitrary points
// (Domain.interp_from_interior_points, _interpolation_methods.py:24-219).
//
// One thread per evaluation point: the barycentric weights of the point
// against its leaf's Chebyshev nodes in each coordinate (the reference's
// formula, exact coincidences replaced by EPS: quadrature/_interpolation.py:
// 184-191), then the tensor-product sum over the leaf's p^d values.
#include <string>

#include "hps_capi_internal.h"

namespace hps {

constexpr int INTERP_MAX_P = 16;

template <typename T>
__global__ __launch_bounds__(256) void interp_points_kernel(int npts, int dim, int p, int nsrc, const double* pts,
                                                            const int* leaf, const double* boxes, const double* cheb,
                                                            int flip_y, const int* grid_to_leaf, const T* F,
                                                            T* out) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= npts) return;
  const int l = leaf[n];
  const double* box = boxes + (long)l * 2 * dim;
  double w[3][INTERP_MAX_P];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    if (d >= dim) break;
    const double lo = box[2 * d], hi = box[2 * d + 1];
    const double x = pts[(long)n * dim + d];
    const bool rev = dim == 2 && d == 1 && flip_y;
    double t[INTERP_MAX_P];
#pragma unroll
    for (int a = 0; a < INTERP_MAX_P; ++a) {
      const int src = rev ? p - 1 - a : a;
      t[a] = a < p ? 0.5 * (hi - lo) * cheb[src < p ? src : 0] + 0.5 * (hi + lo) : 0.0;
    }
    double norm = 0.0;
    double bw[INTERP_MAX_P], dist[INTERP_MAX_P];
#pragma unroll
    for (int a = 0; a < INTERP_MAX_P; ++a) {
      double prod = 1.0;
#pragma unroll
      for (int k = 0; k < INTERP_MAX_P; ++k)
        if (k < p && k != a) prod *= t[k] - t[a];
      bw[a] = prod;
      double dd = x - t[a];
      dist[a] = dd == 0.0 ? 2.220446049250313e-16 : dd;
      if (a < p) norm += 1.0 / (bw[a] * dist[a]);
    }
#pragma unroll
    for (int a = 0; a < INTERP_MAX_P; ++a) w[d][a] = a < p ? 1.0 / (dist[a] * bw[a] * norm) : 0.0;
  }
  const int np = dim == 2 ? p * p : p * p * p;
  const T* Fl = F + (long)l * np * nsrc;
  for (int s = 0; s < nsrc; ++s) {
    T acc = s_zero<T>();
    for (int m = 0; m < np; ++m) {
      const int a = dim == 2 ? m / p : m / (p * p);
      const int b = dim == 2 ? m % p : (m / p) % p;
      double wt = 0.0;
      // runtime indices into the weight tables: read them through a switch-free
      // select over the unrolled entries so the tables stay in registers
#pragma unroll
      for (int i = 0; i < INTERP_MAX_P; ++i)
        if (i == a) wt = w[0][i];
      double wb = 0.0;
#pragma unroll
      for (int i = 0; i < INTERP_MAX_P; ++i)
        if (i == b) wb = w[1][i];
      wt *= wb;
      if (dim == 3) {
        const int c = m % p;
        double wc = 0.0;
#pragma unroll
        for (int i = 0; i < INTERP_MAX_P; ++i)
          if (i == c) wc = w[2][i];
        wt *= wc;
      }
      acc = s_add(acc, s_scale(Fl[(long)grid_to_leaf[m] * nsrc + s], wt));
    }
    out[(long)n * nsrc + s] = acc;
  }
}

template <typename T>
static int interp_points_t(hps_handle_t h, int npts, int dim, int p, int nsrc, const double* pts, const int* leaf,
                           const double* boxes, const double* cheb, int flip_y, const int* grid_to_leaf,
                           const void* F, void* out) {
  hipStream_t s = h->stream;
  HPS_PROF("interp_points", T, 0.0, sizeof(T) * (double)npts * nsrc * 2);
  hipLaunchKernelGGL((interp_points_kernel<T>), dim3(ceil_div(npts, 256)), dim3(256), 0, s, npts, dim, p, nsrc, pts,
                     leaf, boxes, cheb, flip_y, grid_to_leaf, (const T*)F, (T*)out);
  HPS_CHECK_HIP(h, hipGetLastError());
  return HPS_OK;
}

}  // namespace hps

using namespace hps;

int hps_interp_points(hps_handle_t h, int dtype, int npts, int dim, int p, int nsrc, const double* pts,
                      const int* leaf, const double* boxes, const double* cheb, int flip_y, const int* grid_to_leaf,
                      const void* F, void* out) {
  if (!h) return HPS_ERR_ARG;
  HPS_ARG(h, dim == 2 || dim == 3, "dim must be 2 or 3");
  HPS_ARG(h, p >= 2 && p <= INTERP_MAX_P, "p must be in [2, 16]");
  HPS_ARG(h, npts >= 0 && nsrc > 0, "bad sizes");
  HPS_ARG(h, dtype == HPS_F64 || dtype == HPS_C128, "interpolation is FP64 / complex128");
  if (npts == 0) return HPS_OK;
  HPS_ARG(h, pts && leaf && boxes && cheb && grid_to_leaf && F && out, "null pointer");
  if (dtype == HPS_F64)
    return interp_points_t<double>(h, npts, dim, p, nsrc, pts, leaf, boxes, cheb, flip_y, grid_to_leaf, F, out);
  return interp_points_t<c128>(h, npts, dim, p, nsrc, pts, leaf, boxes, cheb, flip_y, grid_to_leaf, F, out);
}
