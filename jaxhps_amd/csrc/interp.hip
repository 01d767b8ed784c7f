// Interpolation from the HPS leaf grids to the points of a rectangular grid
// (Domain.interp_from_interior_points, _interpolation_methods.py:24-219).
//
// One thread per evaluation point: the point's leaf (the first in leaf order
// whose closed box holds it), the barycentric weights of the point against
// that leaf's Chebyshev nodes in each coordinate (the reference's formula,
// exact coincidences replaced by EPS: quadrature/_interpolation.py:184-191),
// then the tensor-product sum over the leaf's p^d values.
#include <string>

#include "hps_capi_internal.h"

namespace hps {

constexpr int INTERP_MAX_P = 16;

// Barycentric weights of point x against p Chebyshev nodes mapped to
// [lo, hi] (reversed when rev).  bw: the barycentric weights of the nodes on
// [-1, 1]; an affine map to [lo, hi] scales all of them by one positive
// constant, which the normalised formula cancels.
__device__ __forceinline__ void bary_weights(double x, double lo, double hi, bool rev, int p, const double* cheb,
                                             const double* bw, double* w) {
  double dist[INTERP_MAX_P];
  double norm = 0.0;
#pragma unroll
  for (int a = 0; a < INTERP_MAX_P; ++a) {
    dist[a] = 1.0;
    if (a < p) {
      const int src = rev ? p - 1 - a : a;
      const double dd = x - (0.5 * (hi - lo) * cheb[src] + 0.5 * (hi + lo));
      dist[a] = dd == 0.0 ? 2.220446049250313e-16 : dd;
      norm += 1.0 / (bw[src] * dist[a]);
    }
  }
#pragma unroll
  for (int a = 0; a < INTERP_MAX_P; ++a) w[a] = a < p ? 1.0 / (dist[a] * bw[rev ? p - 1 - a : a] * norm) : 0.0;
}

// Value at point x of leaf l's interpolant for each of nsrc sources; g2l is
// grid_to_leaf (p^dim entries) staged in LDS.
template <typename T>
__device__ __forceinline__ void eval_point(int dim, int p, int nsrc, const double* x, int l, const double* boxes,
                                           const double* cheb, const double* bw, int flip_y, const int* g2l,
                                           const T* F, T* out) {
  const double* box = boxes + (long)l * 2 * dim;
  double w0[INTERP_MAX_P], w1[INTERP_MAX_P], w2[INTERP_MAX_P];
  bary_weights(x[0], box[0], box[1], false, p, cheb, bw, w0);
  bary_weights(x[1], box[2], box[3], dim == 2 && flip_y, p, cheb, bw, w1);
  if (dim == 3) bary_weights(x[2], box[4], box[5], false, p, cheb, bw, w2);
  const int np = dim == 2 ? p * p : p * p * p;
  const T* Fl = F + (long)l * np * nsrc;
  for (int s = 0; s < nsrc; ++s) {
    T acc = s_zero<T>();
    if (dim == 2) {
#pragma unroll
      for (int a = 0; a < INTERP_MAX_P; ++a) {
        if (a < p) {
          T row = s_zero<T>();
#pragma unroll
          for (int b = 0; b < INTERP_MAX_P; ++b)
            if (b < p) row = s_add(row, s_scale(Fl[(long)g2l[a * p + b] * nsrc + s], w1[b]));
          acc = s_add(acc, s_scale(row, w0[a]));
        }
      }
    } else {
      for (int a = 0; a < p; ++a) {
        // runtime index into w0: a select over the unrolled entries keeps it
        // in registers
        double wa = 0.0;
#pragma unroll
        for (int i = 0; i < INTERP_MAX_P; ++i)
          if (i == a) wa = w0[i];
        T plane = s_zero<T>();
#pragma unroll
        for (int b = 0; b < INTERP_MAX_P; ++b) {
          if (b < p) {
            T row = s_zero<T>();
#pragma unroll
            for (int c = 0; c < INTERP_MAX_P; ++c)
              if (c < p) row = s_add(row, s_scale(Fl[(long)g2l[(a * p + b) * p + c] * nsrc + s], w2[c]));
            plane = s_add(plane, s_scale(row, w1[b]));
          }
        }
        acc = s_add(acc, s_scale(plane, wa));
      }
    }
    out[s] = acc;
  }
}

// Points of meshgrid(xs, ys[, zs]) in its default ("xy") order: flat index
// (iy * nx + ix) * nz + iz.  Each point's leaf is the first in leaf order
// whose closed box holds it, found among the (up to 2^dim) cells of the
// uniform n_side^dim grid of leaves that can hold it; leaf_of_cell maps cell
// (i_x, i_y[, i_z]) (x slowest) to its leaf.  Points outside every box take
// leaf 0.
template <typename T>
__global__ __launch_bounds__(256) void interp_grid_kernel(int nx, int ny, int nz, int dim, int p, int nsrc,
                                                          const double* xs, const double* ys, const double* zs,
                                                          int n_side, const double* root, const int* leaf_of_cell,
                                                          const double* boxes, const double* cheb, int flip_y,
                                                          const int* grid_to_leaf, const T* F, T* out) {
  __shared__ double s_bw[INTERP_MAX_P];
  __shared__ int s_g2l[INTERP_MAX_P * INTERP_MAX_P * INTERP_MAX_P];
  const int np = dim == 2 ? p * p : p * p * p;
  for (int i = threadIdx.x; i < np; i += blockDim.x) s_g2l[i] = grid_to_leaf[i];
  if (threadIdx.x < p) {
    double prod = 1.0;
    for (int k = 0; k < p; ++k)
      if (k != (int)threadIdx.x) prod *= cheb[k] - cheb[threadIdx.x];
    s_bw[threadIdx.x] = prod;
  }
  __syncthreads();
  const long npts = (long)nx * ny * nz;
  const long n = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (n >= npts) return;
  const int iz = (int)(n % nz);
  const long r = n / nz;
  const int ix = (int)(r % nx), iy = (int)(r / nx);
  double x[3] = {xs[ix], ys[iy], dim == 3 ? zs[iz] : 0.0};
  int base[3] = {0, 0, 0};
  for (int d = 0; d < dim; ++d) {
    const double lo = root[2 * d], hi = root[2 * d + 1];
    double v = floor((x[d] - lo) / ((hi - lo) / n_side));
    if (!(v >= -1.0)) v = -1.0;  // also NaN
    if (v > n_side) v = n_side;
    base[d] = (int)v;
  }
  int best = 0x7fffffff;
  for (int sh = 0; sh < (1 << dim); ++sh) {
    long cell = 0;
    bool ok = true;
    for (int d = 0; d < dim; ++d) {
      const int c = base[d] - ((sh >> (dim - 1 - d)) & 1);
      if (c < 0 || c >= n_side) ok = false;
      cell = cell * n_side + (c < 0 ? 0 : (c >= n_side ? n_side - 1 : c));
    }
    if (!ok) continue;
    const int l = leaf_of_cell[cell];
    const double* bx = boxes + (long)l * 2 * dim;
    for (int d = 0; d < dim; ++d)
      if (!(x[d] >= bx[2 * d] && x[d] <= bx[2 * d + 1])) ok = false;
    if (ok && l < best) best = l;
  }
  if (best == 0x7fffffff) best = 0;
  eval_point<T>(dim, p, nsrc, x, best, boxes, cheb, s_bw, flip_y, s_g2l, F, out + n * nsrc);
}

template <typename T>
static int interp_grid_t(hps_handle_t h, int nx, int ny, int nz, int dim, int p, int nsrc, const double* xs,
                         const double* ys, const double* zs, int n_side, const double* root, const int* leaf_of_cell,
                         const double* boxes, const double* cheb, int flip_y, const int* grid_to_leaf, const void* F,
                         void* out) {
  hipStream_t s = h->stream;
  const long npts = (long)nx * ny * nz;
  // per point: the weights (about 8 p flops per axis) and the p^dim
  // multiply-adds per source; bytes: every leaf sample read once, the
  // values written once
  const double np = dim == 2 ? (double)p * p : (double)p * p * p;
  const double n_leaves = dim == 2 ? (double)n_side * n_side : (double)n_side * n_side * n_side;
  const double flops = (double)npts * (8.0 * p * dim + 2.0 * np * nsrc * (sizeof(T) == 16 ? 2.0 : 1.0));
  HPS_PROF("interp_grid", T, flops, sizeof(T) * ((double)npts * nsrc + n_leaves * np * nsrc));
  hipLaunchKernelGGL((interp_grid_kernel<T>), dim3(ceil_div(npts, 256)), dim3(256), 0, s, nx, ny, nz, dim, p, nsrc,
                     xs, ys, zs, n_side, root, leaf_of_cell, boxes, cheb, flip_y, grid_to_leaf, (const T*)F, (T*)out);
  HPS_CHECK_HIP(h, hipGetLastError());
  return HPS_OK;
}

}  // namespace hps

using namespace hps;

int hps_interp_grid(hps_handle_t h, int dtype, int nx, int ny, int nz, int dim, int p, int nsrc, const double* xs,
                    const double* ys, const double* zs, int n_side, const double* root, const int* leaf_of_cell,
                    const double* boxes, const double* cheb, int flip_y, const int* grid_to_leaf, const void* F,
                    void* out) {
  if (!h) return HPS_ERR_ARG;
  HPS_ARG(h, dim == 2 || dim == 3, "dim must be 2 or 3");
  HPS_ARG(h, p >= 2 && p <= INTERP_MAX_P, "p must be in [2, 16]");
  HPS_ARG(h, nx >= 0 && ny >= 0 && nsrc > 0 && n_side >= 1, "bad sizes");
  HPS_ARG(h, dim == 3 ? nz >= 0 : nz == 1, "nz must be 1 in 2D");
  HPS_ARG(h, dtype == HPS_F64 || dtype == HPS_C128, "interpolation is FP64 / complex128");
  HPS_ARG(h, (double)nx * ny * nz < 2147483647.0 * 256.0, "too many points");
  if ((long)nx * ny * nz == 0) return HPS_OK;
  HPS_ARG(h, xs && ys && (dim == 2 || zs) && root && leaf_of_cell && boxes && cheb && grid_to_leaf && F && out,
          "null pointer");
  if (dtype == HPS_F64)
    return interp_grid_t<double>(h, nx, ny, nz, dim, p, nsrc, xs, ys, zs, n_side, root, leaf_of_cell, boxes, cheb,
                                 flip_y, grid_to_leaf, F, out);
  return interp_grid_t<c128>(h, nx, ny, nz, dim, p, nsrc, xs, ys, zs, n_side, root, leaf_of_cell, boxes, cheb,
                             flip_y, grid_to_leaf, F, out);
}
