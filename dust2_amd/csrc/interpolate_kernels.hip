// Batched evaluation of an interpolator at many times: one thread per
// (time, series) pair.  Backs dust2_amd_interpolate, the C-ABI check of the
// device interpolation used by models with time-varying inputs.
#include <hip/hip_runtime.h>

#include "engine.hpp"
#include "interpolate.hpp"

namespace dust2_amd {

namespace {

// out[q * n_series + j] = series j at z[q]; err[q] = the first error of
// time z[q] (INTERP_BEFORE / INTERP_AFTER), 0 when it evaluated.
__global__ void __launch_bounds__(256)
    interpolate_kernel(int kind, const double* __restrict__ t, const double* __restrict__ y,
                       const double* __restrict__ k, int n, int n_series,
                       const double* __restrict__ z, int n_z, double* __restrict__ out,
                       int* __restrict__ err) {
  const long idx = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (idx >= static_cast<long>(n_z) * n_series) return;
  const int q = static_cast<int>(idx / n_series);
  const int j = static_cast<int>(idx - static_cast<long>(q) * n_series);
  int e = INTERP_OK;
  out[idx] = interp_eval(kind, t, y, k, n, n_series, j, z[q], e);
  if (j == 0) err[q] = e;
}

}  // namespace

hipError_t launch_interpolate(int kind, const double* t, const double* y, const double* k, int n,
                              int n_series, const double* z, int n_z, double* out, int* err,
                              hipStream_t stream) {
  const long total = static_cast<long>(n_z) * n_series;
  if (total == 0) return hipSuccess;
  const int block = 256;
  const dim3 grid(static_cast<unsigned>((total + block - 1) / block));
  hipLaunchKernelGGL(interpolate_kernel, grid, dim3(block), 0, stream, kind, t, y, k, n,
                     n_series, z, n_z, out, err);
  return hipGetLastError();
}

}  // namespace dust2_amd
