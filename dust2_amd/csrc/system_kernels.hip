// System-level kernels with particle state resident in HBM between calls:
// reorder for dust_system_* style use (dust2/discrete/system.hpp:169-181;
// run_to_time and compare_data are in discrete_kernels.hip), plus the RNG
// stream seeding and layout kernels and small kernels that expose the
// samplers and the resampler to the tests.
//
// Layouts: state real [G][S][P] (one contiguous SoA block per group, so a
// wavefront's loads of one component are coalesced); particle streams
// [G][4][P]; raw streams as exported to callers are AoS [n][4] words, the
// reference's byte layout (dust2/r/helpers.hpp:448-453).
#include <hip/hip_runtime.h>

#include "block_resample.hpp"
#include "device_random.hpp"
#include "engine.hpp"

namespace dust2_amd {

namespace {

template <typename real>
__global__ void __launch_bounds__(256)
    reorder_kernel(const real* __restrict__ in, real* __restrict__ out,
                   const uint32_t* __restrict__ idx, int G, int S, int P) {
  const int g = blockIdx.y;
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P || g >= G) return;
  const size_t base = static_cast<size_t>(g) * S * P;
  const uint32_t from = idx[static_cast<size_t>(g) * P + p];
  for (int c = 0; c < S; ++c) out[base + c * P + p] = in[base + c * P + from];
}

// stream n = jump^n(s0): apply J_i (column-major GF(2) 256x256 matrix,
// column j = J_i e_j as 4 words) for every set bit i of n.
__global__ void __launch_bounds__(256)
    seed_streams_kernel(rng_words s0, const uint64_t* __restrict__ mats, int n_mats,
                        uint64_t n_streams, uint64_t* __restrict__ out, uint64_t soa_stride) {
  const uint64_t n = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (n >= n_streams) return;
  uint64_t s[4] = {s0.w[0], s0.w[1], s0.w[2], s0.w[3]};
  for (int i = 0; i < n_mats; ++i) {
    if (((n >> i) & 1u) == 0) continue;
    const uint64_t* m = mats + static_cast<size_t>(i) * 256 * 4;
    uint64_t acc[4] = {0, 0, 0, 0};
    for (int w = 0; w < 4; ++w) {
      const uint64_t word = s[w];
      for (int b = 0; b < 64; ++b) {
        const uint64_t mask = 0 - ((word >> b) & 1u);
        const uint64_t* col = m + (w * 64 + b) * 4;
        acc[0] ^= col[0] & mask;
        acc[1] ^= col[1] & mask;
        acc[2] ^= col[2] & mask;
        acc[3] ^= col[3] & mask;
      }
    }
    s[0] = acc[0];
    s[1] = acc[1];
    s[2] = acc[2];
    s[3] = acc[3];
  }
  if (soa_stride == 0) {
    out[4 * n + 0] = s[0];
    out[4 * n + 1] = s[1];
    out[4 * n + 2] = s[2];
    out[4 * n + 3] = s[3];
  } else {
    out[n] = s[0];
    out[soa_stride + n] = s[1];
    out[2 * soa_stride + n] = s[2];
    out[3 * soa_stride + n] = s[3];
  }
}

// aos: [G][per_group][4]; soa: [G][4][stride], stride >= per_group
__global__ void rng_aos_to_soa_kernel(const uint64_t* __restrict__ aos, uint64_t* __restrict__ soa,
                                      int G, int per_group, int stride) {
  const int g = blockIdx.y;
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= per_group || g >= G) return;
  const uint64_t* src = aos + (static_cast<size_t>(g) * per_group + j) * 4;
  uint64_t* dst = soa + static_cast<size_t>(g) * 4 * stride + j;
  for (int w = 0; w < 4; ++w) dst[w * stride] = src[w];
}

__global__ void rng_soa_to_aos_kernel(const uint64_t* __restrict__ soa, uint64_t* __restrict__ aos,
                                      int G, int per_group, int stride) {
  const int g = blockIdx.y;
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= per_group || g >= G) return;
  const uint64_t* src = soa + static_cast<size_t>(g) * 4 * stride + j;
  uint64_t* dst = aos + (static_cast<size_t>(g) * per_group + j) * 4;
  for (int w = 0; w < 4; ++w) dst[w] = src[w * stride];
}

// One workgroup of 256 threads resamples n <= 256*8 weights with the same
// code the filter kernels use.
__global__ void __launch_bounds__(256)
    resample_test_kernel(const double* w_in, int n, double u, uint32_t* idx_out, double* ll) {
  __shared__ double smem[kScratchDoubles + 256 * 8];
  double w[8];
  int idx[8];
  const int p0 = threadIdx.x * 8;
#pragma unroll
  for (int k = 0; k < 8; ++k) w[k] = (p0 + k < n) ? w_in[p0 + k] : 0.0;
  const resample_result r = block_resample<256, 8>(w, idx, n, smem + kScratchDoubles, smem,
                                                   [&]() { return u; });
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    if (p0 + k < n) idx_out[p0 + k] = static_cast<uint32_t>(idx[k]);
  }
  if (threadIdx.x == 0) {
    ll[0] = r.ll_increment;
    ll[1] = r.resampled ? 1.0 : 0.0;
    ll[2] = r.n_fallback;
  }
}

// n_draws draws per stream; a, b per stream.  out: [stream][draw].
template <typename real>
__global__ void sample_test_kernel(int kind, uint64_t* rng_aos, int n_streams, int n_draws,
                                   const double* a, const double* b, int det, double* out,
                                   uint32_t* err_out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_streams) return;
  rng256 r{rng_aos[4 * i], rng_aos[4 * i + 1], rng_aos[4 * i + 2], rng_aos[4 * i + 3]};
  const real ai = static_cast<real>(a[i]);
  const real bi = static_cast<real>(b[i]);
  uint32_t err = 0;
  for (int d = 0; d < n_draws; ++d) {
    real v;
    switch (kind) {
      case 0: v = binomial<real>(r, ai, bi, det != 0, err); break;
      case 1: v = exponential_rate<real>(r, ai, det != 0); break;
      case 2: v = normal<real>(r, ai, bi, det != 0); break;
      case 3: v = poisson<real>(r, ai, det != 0); break;
      default: v = random_real<real>(r); break;
    }
    out[static_cast<size_t>(i) * n_draws + d] = static_cast<double>(v);
    if (err) break;
  }
  err_out[i] = err;
  rng_aos[4 * i] = r.s0;
  rng_aos[4 * i + 1] = r.s1;
  rng_aos[4 * i + 2] = r.s2;
  rng_aos[4 * i + 3] = r.s3;
}

dim3 grid_gp(int G, int P, int block) {
  return dim3(static_cast<unsigned>((P + block - 1) / block), static_cast<unsigned>(G));
}

}  // namespace

hipError_t launch_reorder(const void* state_in, void* state_out, const uint32_t* idx, int G,
                          int S, int P, bool fp64, hipStream_t stream) {
  if (G == 0 || P == 0) return hipSuccess;
  if (fp64) {
    hipLaunchKernelGGL(reorder_kernel<double>, grid_gp(G, P, 256), dim3(256), 0, stream,
                       static_cast<const double*>(state_in), static_cast<double*>(state_out),
                       idx, G, S, P);
  } else {
    hipLaunchKernelGGL(reorder_kernel<float>, grid_gp(G, P, 256), dim3(256), 0, stream,
                       static_cast<const float*>(state_in), static_cast<float*>(state_out), idx,
                       G, S, P);
  }
  return hipGetLastError();
}

hipError_t launch_seed_streams(const uint64_t s0[4], const uint64_t* d_mats, int n_mats,
                               uint64_t n_streams, uint64_t* d_out, uint64_t soa_stride,
                               hipStream_t stream) {
  if (n_streams == 0) return hipSuccess;
  const unsigned blocks = static_cast<unsigned>((n_streams + 255) / 256);
  const rng_words w{{s0[0], s0[1], s0[2], s0[3]}};
  hipLaunchKernelGGL(seed_streams_kernel, dim3(blocks), dim3(256), 0, stream, w, d_mats,
                     n_mats, n_streams, d_out, soa_stride);
  return hipGetLastError();
}

hipError_t launch_rng_aos_to_soa(const uint64_t* aos, uint64_t* soa, int G, int per_group,
                                 int stride, hipStream_t stream) {
  if (G == 0 || per_group == 0) return hipSuccess;
  hipLaunchKernelGGL(rng_aos_to_soa_kernel, grid_gp(G, per_group, 256), dim3(256), 0, stream,
                     aos, soa, G, per_group, stride);
  return hipGetLastError();
}

hipError_t launch_rng_soa_to_aos(const uint64_t* soa, uint64_t* aos, int G, int per_group,
                                 int stride, hipStream_t stream) {
  if (G == 0 || per_group == 0) return hipSuccess;
  hipLaunchKernelGGL(rng_soa_to_aos_kernel, grid_gp(G, per_group, 256), dim3(256), 0, stream,
                     soa, aos, G, per_group, stride);
  return hipGetLastError();
}

hipError_t launch_resample_test(const double* w, int n, double u, uint32_t* idx,
                                double* ll_increment, hipStream_t stream) {
  if (n < 1 || n > 256 * 8) return hipErrorInvalidValue;
  hipLaunchKernelGGL(resample_test_kernel, dim3(1), dim3(256), 0, stream, w, n, u, idx,
                     ll_increment);
  return hipGetLastError();
}

hipError_t launch_sample_test(int kind, bool fp64, uint64_t* rng_aos, int n_streams,
                              int n_draws, const double* a, const double* b, int det,
                              double* out, uint32_t* err, hipStream_t stream) {
  if (n_streams == 0) return hipSuccess;
  const unsigned blocks = static_cast<unsigned>((n_streams + 255) / 256);
  if (fp64) {
    hipLaunchKernelGGL(sample_test_kernel<double>, dim3(blocks), dim3(256), 0, stream, kind,
                       rng_aos, n_streams, n_draws, a, b, det, out, err);
  } else {
    hipLaunchKernelGGL(sample_test_kernel<float>, dim3(blocks), dim3(256), 0, stream, kind,
                       rng_aos, n_streams, n_draws, a, b, det, out, err);
  }
  return hipGetLastError();
}

}  // namespace dust2_amd
