// Vector-ALU issue-rate calibration for the measurement of the sampler
// kernels, whose time goes to vector instructions rather than to HBM: a
// kernel of independent dependency chains (eight FP32 or four FP64 fused
// multiply-adds per iteration, so that issue rather than latency limits a
// full SIMD) run at a chosen number of wavefronts per SIMD.  The measured
// wave-instructions per SIMD per nanosecond is the peak against which the
// kernels' own VALU instruction counts (rocprofv3 SQ_INSTS_VALU) are set.
#include <hip/hip_runtime.h>

#include "engine.hpp"

namespace dust2_amd {

namespace {

constexpr int kProbeBlock = 256;  // four wavefronts

template <bool FP64>
__global__ void __launch_bounds__(kProbeBlock) valu_probe_kernel(double* out, int iters,
                                                                 float a) {
  float x[8];
  double y[4];
#pragma unroll
  for (int i = 0; i < 8; ++i) x[i] = threadIdx.x * 1e-3f + i;
#pragma unroll
  for (int i = 0; i < 4; ++i) y[i] = threadIdx.x * 1e-3 + i;
  for (int it = 0; it < iters; ++it) {
    if constexpr (FP64) {
#pragma unroll
      for (int i = 0; i < 4; ++i) y[i] = __builtin_fma(y[i], static_cast<double>(a), 0.5);
    } else {
#pragma unroll
      for (int i = 0; i < 8; ++i) x[i] = __builtin_fmaf(x[i], a, 0.5f);
    }
  }
  double s = 0;
#pragma unroll
  for (int i = 0; i < 8; ++i) s += x[i];
#pragma unroll
  for (int i = 0; i < 4; ++i) s += y[i];
  out[blockIdx.x * kProbeBlock + threadIdx.x] = s;
}

}  // namespace

int valu_probe_fmas_per_iteration(bool fp64) { return fp64 ? 4 : 8; }

hipError_t launch_valu_probe(bool fp64, int n_blocks, int iters, double* out,
                             hipStream_t stream) {
  if (fp64) {
    hipLaunchKernelGGL(valu_probe_kernel<true>, dim3(n_blocks), dim3(kProbeBlock), 0, stream,
                       out, iters, 0.999f);
  } else {
    hipLaunchKernelGGL(valu_probe_kernel<false>, dim3(n_blocks), dim3(kProbeBlock), 0, stream,
                       out, iters, 0.999f);
  }
  return hipGetLastError();
}

}  // namespace dust2_amd
