// Interval pipeline for the bootstrap particle filter: particle state lives
// in HBM between data intervals and each interval is two launches,
//
//   advance  (one thread per particle): gather the particle's state from its
//            ancestor slot, run the interval's steps, draw the compare noise
//            and write the Poisson mean lambda;
//   resample (one workgroup per group): FP64 log-weights from lambda,
//            log-sum-exp, the group's log-likelihood and systematic
//            resampling into ancestor indices for the next advance.
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
#include "interpolate.hpp"
#include "models.hpp"

namespace dust2_amd {

namespace {

template <typename real>
__device__ __forceinline__ sir_model::shared<real> sir_shared_from(const double* pars) {
  sir_model::shared<real> sh;
  sh.N = static_cast<real>(pars[0]);
  sh.I0 = static_cast<real>(pars[1]);
  sh.beta = static_cast<real>(pars[2]);
  sh.gamma = static_cast<real>(pars[3]);
  sh.exp_noise = pars[4];
  sh.p_IR = static_cast<real>(sizeof(real) == 8 ? pars[5] : pars[6]);
  return sh;
}

constexpr int kAdvBlock = 256;

template <typename real>
__global__ void __launch_bounds__(kAdvBlock)
    sir_advance_kernel(const pipe_slot* __restrict__ slots, int t, int parity, uint32_t mode) {
  const pipe_slot& sd = slots[blockIdx.y];
  const int P = static_cast<int>(sd.n_particles);
  const int p = blockIdx.x * kAdvBlock + threadIdx.x;
  if (p >= P) return;
  if (!pipe_slot_interval(sd, t, parity, mode)) return;
  // the reference stops at the first interval with errors
  // (dust2/discrete/system.hpp:87): later intervals of the run do nothing
  if ((mode & PIPE_STEP) && __atomic_load_n(sd.error_count, __ATOMIC_RELAXED) != 0) return;
  const bool det = (sd.flags & SLOT_DETERMINISTIC) != 0;
  const sir_model::shared<real> sh = sir_shared_from<real>(sd.pars);
  const real* state_in = static_cast<const real*>(parity ? sd.state_b : sd.state_a);
  real* state_out = static_cast<real*>(parity ? sd.state_a : sd.state_b);
  const int rs = P + 2;

  real y[5];
  if (mode & PIPE_INITIAL) {
    sir_model::initial(sh, y);
  } else {
    const int a = (mode & PIPE_GATHER) ? static_cast<int>(sd.ancestors[p]) : p;
#pragma unroll
    for (int c = 0; c < 5; ++c) y[c] = state_in[c * P + a];
  }
  if ((mode & PIPE_TRAJ_PREV) && sd.traj) {
    // post-reorder state of the previous interval (trajectories_.add)
    real* tr = static_cast<real*>(sd.traj) + static_cast<size_t>(t - 1) * 5 * P;
#pragma unroll
    for (int c = 0; c < 5; ++c) tr[c * P + p] = y[c];
  }
  if (mode & PIPE_STEP) {
    rng256 r{sd.rng[2 + p], sd.rng[rs + 2 + p], sd.rng[2 * rs + 2 + p], sd.rng[3 * rs + 2 + p]};
    const int n_steps = static_cast<int>(sd.sched_steps[t]);
    const uint64_t zero_mask = sd.sched_zero[t];
    const real dt_r = static_cast<real>(sd.dt);
    bool failed = false;
    for (int s = 0; s < n_steps; ++s) {
      // steps beyond 64 are zeroed by the host-computed rule replicated here
      const bool zero = s < 64 ? ((zero_mask >> s) & 1u) != 0
                               : fmod(sd.sched_time[t] + s * sd.dt, sir_model::zero_period) == 0;
      if (zero) y[sir_model::zero_first] = 0;
      uint32_t err = 0;
      real ea = 0, eb = 0;
      sir_model::shared<real> sh_s = sh;
      if (sd.beta_n > 0) {
        // beta at the step's start time, as the step's update sees it
        const double time_s = sd.sched_time[t] + s * sd.dt;
        int ie = INTERP_OK;
        const double b = interp_eval(sd.beta_kind, sd.beta_t, sd.beta_y, sd.beta_k, sd.beta_n,
                                     1, 0, time_s, ie);
        if (ie != INTERP_OK) {
          const double knot = ie == INTERP_BEFORE ? sd.beta_t[0] : sd.beta_t[sd.beta_n - 1];
          sd.errors[p] = error_record{ie == INTERP_BEFORE ? ERR_INTERP_BEFORE : ERR_INTERP_AFTER,
                                      static_cast<uint32_t>(t), time_s, knot};
          atomicAdd(sd.error_count, 1u);
          failed = true;
          break;
        }
        sh_s.beta = static_cast<real>(b);
      }
      sir_model::update<real>(dt_r, y, sh_s, r, det, err, ea, eb);
      if (err) {
        sd.errors[p] = error_record{err, static_cast<uint32_t>(t), static_cast<double>(ea),
                                    static_cast<double>(eb)};
        atomicAdd(sd.error_count, 1u);
        failed = true;
        break;
      }
    }
    // compare noise (sir.cpp:111-118): no draw when the datum is missing
    const double x = sd.obs[static_cast<size_t>(t) * sd.obs_stride];
    double lambda = 0;
    if (!isnan(x) && !failed) {
      double noise;
      if (sizeof(real) == 8) {
        noise = exponential_rate<double>(r, sh.exp_noise, det);
      } else {
        noise = static_cast<double>(
            exponential_rate<float>(r, static_cast<float>(sh.exp_noise), det));
      }
      lambda = static_cast<double>(y[4]) + noise;
    }
    sd.lambda[p] = lambda;
    sd.rng[2 + p] = r.s0;
    sd.rng[rs + 2 + p] = r.s1;
    sd.rng[2 * rs + 2 + p] = r.s2;
    sd.rng[3 * rs + 2 + p] = r.s3;
  }
#pragma unroll
  for (int c = 0; c < 5; ++c) state_out[c * P + p] = y[c];
}

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
  const double x = sd.obs[static_cast<size_t>(t) * sd.obs_stride];
  const double lf = sd.lfact[static_cast<size_t>(t) * sd.obs_stride];
  const bool is_ll = (sd.flags & SLOT_WEIGHT_IS_LL) != 0;
  double w[PPT];
#pragma unroll
  for (int k = 0; k < PPT; ++k) {
    w[k] = 0;
    if (p0 + k < P) {
      if (is_ll) {
        w[k] = sd.lambda[p0 + k];
      } else if (!isnan(x)) {
        w[k] = dpois_log(x, sd.lambda[p0 + k], lf);
      }
    }
  }
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
int pipe_resample_ppt(int max_particles) {
  if (max_particles <= 1024) return 1;
  if (max_particles <= 2048) return 2;
  if (max_particles <= 4096) return 4;
  if (max_particles <= 8192) return 8;
  if (max_particles <= 16384) return 16;
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

hipError_t launch_sir_advance(const pipe_slot* d_slots, int n_slots, int max_particles, int t,
                              int parity, uint32_t mode, bool fp64, hipStream_t stream) {
  if (n_slots == 0) return hipSuccess;
  const dim3 grid((max_particles + kAdvBlock - 1) / kAdvBlock, n_slots);
  if (fp64) {
    hipLaunchKernelGGL(sir_advance_kernel<double>, grid, dim3(kAdvBlock), 0, stream, d_slots, t,
                       parity, mode);
  } else {
    hipLaunchKernelGGL(sir_advance_kernel<float>, grid, dim3(kAdvBlock), 0, stream, d_slots, t,
                       parity, mode);
  }
  return hipGetLastError();
}

hipError_t launch_pipe_resample(const pipe_slot* d_slots, int n_slots, int max_particles, int t,
                                hipStream_t stream) {
  if (n_slots == 0) return hipSuccess;
  const size_t lds = (kScratchDoubles + static_cast<size_t>(max_particles)) * sizeof(double);
  switch (pipe_resample_ppt(max_particles)) {
    case 1:
      hipLaunchKernelGGL((resample_kernel<1024, 1>), dim3(n_slots), dim3(1024), lds, stream,
                         d_slots, t);
      break;
    case 2:
      hipLaunchKernelGGL((resample_kernel<1024, 2>), dim3(n_slots), dim3(1024), lds, stream,
                         d_slots, t);
      break;
    case 4:
      hipLaunchKernelGGL((resample_kernel<1024, 4>), dim3(n_slots), dim3(1024), lds, stream,
                         d_slots, t);
      break;
    case 8:
      hipLaunchKernelGGL((resample_kernel<1024, 8>), dim3(n_slots), dim3(1024), lds, stream,
                         d_slots, t);
      break;
    case 16:
      hipLaunchKernelGGL((resample_kernel<1024, 16>), dim3(n_slots), dim3(1024), lds, stream,
                         d_slots, t);
      break;
    default:
      return hipErrorInvalidConfiguration;
  }
  return hipGetLastError();
}

}  // namespace dust2_amd
