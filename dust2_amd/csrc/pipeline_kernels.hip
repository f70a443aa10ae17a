// Interval pipeline for the bootstrap particle filter: particle state lives
// in HBM between data intervals and each interval is two launches,
//
//   advance  (discrete_kernels.hip, one thread per particle): gather the
//            particle's state from its ancestor slot, run the interval's
//            steps and write the FP64 log-weight of compare_data;
//   resample (here, one workgroup per group): log-sum-exp, the group's
//            log-likelihood and systematic resampling into ancestor
//            indices for the next advance.
//
// This is the reference's run_to_time -> compare_data -> scale_log_weights
// -> resample_weight -> reorder cycle (dust2/filter.hpp:136-172) with the
// reorder fused into the next interval's load.  Compared with the
// persistent kernel it holds one particle per thread (low register use,
// high occupancy, any particle count), at the cost of the state round trip
// through HBM every interval.
#include <hip/hip_runtime.h>

#include "block_resample.hpp"
#include "device_random.hpp"
#include "engine.hpp"

namespace dust2_amd {

namespace {

template <int BLOCK, int PPT>
__global__ void __launch_bounds__(BLOCK) resample_kernel(const pipe_slot* __restrict__ slots,
                                                         int t) {
  extern __shared__ double smem[];
  double* scratch = smem;
  double* cum = smem + kScratchDoubles;
  const pipe_slot& sd = slots[blockIdx.x];
  const int P = static_cast<int>(sd.n_particles);
  {
    int parity = 0;
    uint32_t mode = 0;
    if (!pipe_slot_interval(sd, t, parity, mode)) return;
  }
  const int tid = threadIdx.x;
  const int p0 = tid * PPT;
  double w[PPT];
#pragma unroll
  for (int k = 0; k < PPT; ++k) w[k] = p0 + k < P ? sd.lambda[p0 + k] : 0.0;
  const int rs = P + 2;
  rng256 frng{0, 0, 0, 0};
  if (tid == 0) frng = rng256{sd.rng[0], sd.rng[rs], sd.rng[2 * rs], sd.rng[3 * rs]};
  int idx[PPT];
  const resample_result res = block_resample<BLOCK, PPT>(
      w, idx, P, cum, scratch, [&]() { return random_real<double>(frng); });
#pragma unroll
  for (int k = 0; k < PPT; ++k) {
    if (p0 + k < P) {
      sd.ancestors[p0 + k] = static_cast<uint32_t>(idx[k]);
      if (sd.ancestry) sd.ancestry[static_cast<size_t>(t) * P + p0 + k] = idx[k];
    }
  }
  if (tid == 0) {
    sd.ll[0] += res.ll_increment;
    sd.rng[0] = frng.s0;
    sd.rng[rs] = frng.s1;
    sd.rng[2 * rs] = frng.s2;
    sd.rng[3 * rs] = frng.s3;
    if (res.n_fallback) atomicAdd(sd.n_fallback, 1u);
  }
}

// Log-likelihood and error count of every slot into one array (one copy
// back per scheduling tick instead of one per slot).
__global__ void slot_status_kernel(const pipe_slot* __restrict__ slots, int n_slots,
                                   double* out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n_slots) {
    out[2 * i] = slots[i].ll[0];
    out[2 * i + 1] = static_cast<double>(slots[i].error_count[0]);
  }
}

// Zero every slot's log-likelihood (a kernel, not a memset, so the whole
// run can be captured in a graph).
__global__ void reset_ll_kernel(const pipe_slot* __restrict__ slots, int n_slots) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n_slots) slots[i].ll[0] = 0;
}

}  // namespace

// Particles per thread of the 1024-thread resample workgroup; 0 when a
// group is too large for one workgroup.
int pipe_resample_ppt(int max_particles, int block) {
  // widths offered: 1024 threads up to 16 particles each, 512 and 256 up to 8
  const int max_ppt = block == 1024 ? 16 : (block == 512 || block == 256) ? 8 : 0;
  for (int ppt = 1; ppt <= max_ppt; ppt *= 2) {
    if (block * ppt >= max_particles) return ppt;
  }
  return 0;
}

hipError_t launch_slot_status(const pipe_slot* d_slots, int n_slots, double* out,
                              hipStream_t stream) {
  if (n_slots == 0) return hipSuccess;
  hipLaunchKernelGGL(slot_status_kernel, dim3((n_slots + 255) / 256), dim3(256), 0, stream,
                     d_slots, n_slots, out);
  return hipGetLastError();
}

hipError_t launch_pipe_reset(const pipe_slot* d_slots, int n_slots, hipStream_t stream) {
  hipLaunchKernelGGL(reset_ll_kernel, dim3((n_slots + 255) / 256), dim3(256), 0, stream, d_slots,
                     n_slots);
  return hipGetLastError();
}

namespace {

template <int BLOCK, int PPT>
void launch_resample_shape(const pipe_slot* d_slots, int n_slots, size_t lds, int t,
                           hipStream_t stream) {
  hipLaunchKernelGGL((resample_kernel<BLOCK, PPT>), dim3(n_slots), dim3(BLOCK), lds, stream,
                     d_slots, t);
}

template <int BLOCK>
bool launch_resample_ppt(int ppt, const pipe_slot* d_slots, int n_slots, size_t lds, int t,
                         hipStream_t stream) {
  switch (ppt) {
    case 1: launch_resample_shape<BLOCK, 1>(d_slots, n_slots, lds, t, stream); return true;
    case 2: launch_resample_shape<BLOCK, 2>(d_slots, n_slots, lds, t, stream); return true;
    case 4: launch_resample_shape<BLOCK, 4>(d_slots, n_slots, lds, t, stream); return true;
    case 8: launch_resample_shape<BLOCK, 8>(d_slots, n_slots, lds, t, stream); return true;
    case 16:
      if constexpr (BLOCK == 1024) {
        launch_resample_shape<BLOCK, 16>(d_slots, n_slots, lds, t, stream);
        return true;
      }
      return false;
    default: return false;
  }
}

}  // namespace

hipError_t launch_pipe_resample(const pipe_slot* d_slots, int n_slots, int max_particles, int t,
                                hipStream_t stream, int block) {
  if (n_slots == 0) return hipSuccess;
  const size_t lds = (kScratchDoubles + static_cast<size_t>(max_particles)) * sizeof(double);
  const int ppt = pipe_resample_ppt(max_particles, block);
  bool ok = false;
  switch (block) {
    case 1024: ok = launch_resample_ppt<1024>(ppt, d_slots, n_slots, lds, t, stream); break;
    case 512: ok = launch_resample_ppt<512>(ppt, d_slots, n_slots, lds, t, stream); break;
    case 256: ok = launch_resample_ppt<256>(ppt, d_slots, n_slots, lds, t, stream); break;
    default: break;
  }
  if (!ok) return hipErrorInvalidConfiguration;
  return hipGetLastError();
}

}  // namespace dust2_amd
