// Small batched kernels for adaptive (non-uniform) trees, where leaves and
// merges come in different sizes: per-leaf scaling of coefficient fields and
// maps by powers of the leaf's half side, and strided copies that pack
// children of different sizes into the padded blocks a merge level reads.
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

template <typename T>
static int copy_t(hps_handle_t h, int rows, int cols, double alpha, const void* src, long lds, long ss, void* dst,
                  long ldd, long sd, int batch) {
  HPS_CHECK_HIP(h, copy2d<T>(rows, cols, alpha, (const T*)src, lds, ss, (T*)dst, ldd, sd, batch, h->stream));
  return HPS_OK;
}

}  // namespace hps

using namespace hps;

int hps_scale_batched(hps_handle_t h, int dtype, int batch, long n, long stride, const double* scale, void* X) {
  if (!h) return HPS_ERR_ARG;
  HPS_ARG(h, batch >= 0 && n >= 0 && stride >= n, "bad size (stride >= n)");
  if ((long)batch * n == 0) return HPS_OK;
  HPS_ARG(h, scale && X, "null pointer");
  DISPATCH(dtype, scale_t, h, batch, n, stride, scale, X);
}

int hps_copy2d_batched(hps_handle_t h, int dtype, int rows, int cols, double alpha, const void* src, long lds,
                       long ss, void* dst, long ldd, long sd, int batch) {
  if (!h) return HPS_ERR_ARG;
  HPS_ARG(h, rows >= 0 && cols >= 0 && batch >= 0 && lds >= cols && ldd >= cols, "bad size");
  if ((long)rows * cols * batch == 0) return HPS_OK;
  HPS_ARG(h, dst, "null pointer");
  HPS_ARG(h, src != dst, "copy is out of place");
  DISPATCH(dtype, copy_t, h, rows, cols, alpha, src, lds, ss, dst, ldd, sd, batch);
}
