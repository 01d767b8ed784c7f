// Small batched kernels for adaptive (non-uniform) trees, where leaves and
// merges come in different sizes: per-leaf scaling of coefficient fields and
// maps by powers of the leaf's half side, and one-launch block copies that
// pack children of different sizes into the padded blocks a merge level reads.
#include <string>

#include "hps_capi_internal.h"

namespace hps {

// X[b*stride + i] *= scale[b] for i < n
template <typename T>
__global__ __launch_bounds__(256) void scale_batched_kernel(int batch, long n, long stride, const double* scale,
                                                            T* X) {
  const long total = (long)batch * n;
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < total; t += (long)gridDim.x * blockDim.x) {
    const long b = t / n;
    const long i = t - b * n;
    T* x = X + b * stride + i;
    *x = s_scale(*x, scale[b]);
  }
}

template <typename T>
static int scale_t(hps_handle_t h, int batch, long n, long stride, const double* scale, void* X) {
  hipStream_t s = h->stream;
  HPS_PROF("scale_batched", T, 0.0, 2.0 * sizeof(T) * (double)batch * n);
  hipLaunchKernelGGL((scale_batched_kernel<T>), grid_for((long)batch * n, 1), dim3(256), 0, s, batch, n, stride,
                     scale, (T*)X);
  HPS_CHECK_HIP(h, hipGetLastError());
  return HPS_OK;
}

// One launch copying `count` blocks between arbitrary device addresses:
// desc[6*e .. 6*e+5] = (src, dst, rows, cols, lds, ldd) with src/dst byte
// addresses.  blockIdx.y walks the blocks, blockIdx.x the elements.
template <typename T>
__global__ __launch_bounds__(256) void pack_blocks_kernel(int count, const long long* desc) {
  for (int e = blockIdx.y; e < count; e += gridDim.y) {
    const long long* d = desc + 6 * (long)e;
    const T* src = reinterpret_cast<const T*>(d[0]);
    T* dst = reinterpret_cast<T*>(d[1]);
    const long rows = d[2], cols = d[3], lds = d[4], ldd = d[5];
    const long total = rows * cols;
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < total; t += (long)gridDim.x * blockDim.x) {
      const long r = t / cols, c = t - r * cols;
      dst[r * ldd + c] = src[r * lds + c];
    }
  }
}

template <typename T>
static int pack_t(hps_handle_t h, int count, long max_elems, const long long* desc) {
  hipStream_t s = h->stream;
  unsigned gx = ceil_div(max_elems, 256);
  if (gx > 64) gx = 64;
  if (gx == 0) gx = 1;
  unsigned gy = count < 65535 ? count : 65535;
  HPS_PROF("pack_blocks", T, 0.0, 2.0 * sizeof(T) * (double)max_elems * count);
  hipLaunchKernelGGL((pack_blocks_kernel<T>), dim3(gx, gy), dim3(256), 0, s, count, desc);
  HPS_CHECK_HIP(h, hipGetLastError());
  return HPS_OK;
}

}  // namespace hps

using namespace hps;

int hps_pack_blocks(hps_handle_t h, int dtype, int count, long max_elems, const long long* desc) {
  if (!h) return HPS_ERR_ARG;
  HPS_ARG(h, count >= 0 && max_elems >= 0, "bad size");
  if (count == 0 || max_elems == 0) return HPS_OK;
  HPS_ARG(h, desc, "null pointer");
  DISPATCH(dtype, pack_t, h, count, max_elems, desc);
}

int hps_scale_batched(hps_handle_t h, int dtype, int batch, long n, long stride, const double* scale, void* X) {
  if (!h) return HPS_ERR_ARG;
  HPS_ARG(h, batch >= 0 && n >= 0 && stride >= n, "bad size (stride >= n)");
  if ((long)batch * n == 0) return HPS_OK;
  HPS_ARG(h, scale && X, "null pointer");
  DISPATCH(dtype, scale_t, h, batch, n, stride, scale, X);
}
