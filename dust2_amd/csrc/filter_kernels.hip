// Bootstrap particle filter kernels for the stochastic SIR model.
//
// Persistent design: one workgroup owns one parameter group for a whole
// run (or a chunk of data intervals).  Particle state and the particles'
// xoshiro256+ streams stay in registers across every interval; the only
// memory traffic inside the time loop is the observation for the interval
// (one scalar load) and the LDS used for the group's weights and the
// ancestor gather.  This replaces the reference's per-interval sequence
// run_to_time -> compare_data -> scale_log_weights -> resample_weight ->
// reorder (dust2/filter.hpp:136-172; dust2/discrete/system.hpp:61-92,
// 169-181, 252-287) with a single launch.
//
// Thread t owns particles t*PPT .. t*PPT+PPT-1 (their slots, and the slots'
// RNG streams: streams stay with slots on reorder, vignettes/design.Rmd:133).
#include <hip/hip_runtime.h>

#include "block_resample.hpp"
#include "device_random.hpp"
#include "engine.hpp"
#include "models.hpp"

namespace dust2_amd {

namespace {

template <typename real>
struct sir_particle {
  real y[5];
};

template <typename real>
__device__ __forceinline__ sir_model::shared<real> sir_shared(const double* pars) {
  sir_model::shared<real> sh;
  sh.N = static_cast<real>(pars[0]);
  sh.I0 = static_cast<real>(pars[1]);
  sh.beta = static_cast<real>(pars[2]);
  sh.gamma = static_cast<real>(pars[3]);
  sh.exp_noise = pars[4];
  sh.p_IR = static_cast<real>(sizeof(real) == 8 ? pars[5] : pars[6]);
  return sh;
}

template <typename real, int BLOCK, int PPT>
__global__ void __launch_bounds__(BLOCK) sir_filter_kernel(const filter_slot* __restrict__ slots) {
  extern __shared__ double smem[];
  double* scratch = smem;
  double* cum = smem + kScratchDoubles;

  const filter_slot& sd = slots[blockIdx.x];
  const int P = static_cast<int>(sd.n_particles);
  const int tid = threadIdx.x;
  const int p0 = tid * PPT;
  const uint32_t flags = sd.flags;
  const bool det = (flags & SLOT_DETERMINISTIC) != 0;
  const sir_model::shared<real> sh = sir_shared<real>(sd.pars);
  const double dt = sd.dt;
  const real dt_r = static_cast<real>(dt);
  real* state = static_cast<real*>(sd.state);
  const int rs = P + 2;  // rng stride: 2 filter streams then P particles

  sir_particle<real> part[PPT];
  rng256 rng[PPT];
  uint32_t fail_mask = 0;
#pragma unroll
  for (int k = 0; k < PPT; ++k) {
    const int p = p0 + k;
    if (p < P) {
      if (flags & SLOT_SET_INITIAL) {
        sir_model::initial(sh, part[k].y);
      } else {
#pragma unroll
        for (int c = 0; c < 5; ++c) part[k].y[c] = state[c * P + p];
      }
      rng[k].s0 = sd.rng[0 * rs + 2 + p];
      rng[k].s1 = sd.rng[1 * rs + 2 + p];
      rng[k].s2 = sd.rng[2 * rs + 2 + p];
      rng[k].s3 = sd.rng[3 * rs + 2 + p];
    } else {
#pragma unroll
      for (int c = 0; c < 5; ++c) part[k].y[c] = 0;
      rng[k] = rng256{0, 0, 0, 0};
    }
  }
  rng256 frng{0, 0, 0, 0};
  double ll = 0;
  if (tid == 0) {
    frng.s0 = sd.rng[0 * rs];
    frng.s1 = sd.rng[1 * rs];
    frng.s2 = sd.rng[2 * rs];
    frng.s3 = sd.rng[3 * rs];
    ll = sd.ll[0];
  }
  double time = sd.time[0];
  uint32_t fallbacks = 0;

  for (uint32_t t = sd.t_begin; t < sd.t_end; ++t) {
    const double t_target = sd.times[t];
    const int n_steps = static_cast<int>(round(fmax(0.0, t_target - time) / dt));
    const double* obs = sd.obs + static_cast<size_t>(t) * sd.obs_stride;
    const double* lfact = sd.lfact + static_cast<size_t>(t) * sd.obs_stride;
    // zero_every schedule of this interval, computed once: bit s set when
    // fmod(time + s*dt, period) == 0 (dust2/discrete/system.hpp:411-419)
    const bool zero_generic = n_steps > 64;
    uint64_t zero_mask = 0;
    if (!zero_generic) {
      for (int s = 0; s < n_steps; ++s) {
        if (fmod(time + s * dt, sir_model::zero_period) == 0) zero_mask |= uint64_t(1) << s;
      }
    }

    // the thread's particles advance step by step together, so the
    // independent sampler chains of its PPT particles interleave
    double w[PPT];
    bool live[PPT];
#pragma unroll
    for (int k = 0; k < PPT; ++k) live[k] = (p0 + k < P) && !((fail_mask >> k) & 1u);
    for (int s = 0; s < n_steps; ++s) {
      const bool zero = zero_generic ? fmod(time + s * dt, sir_model::zero_period) == 0
                                     : ((zero_mask >> s) & 1u) != 0;
#pragma unroll
      for (int k = 0; k < PPT; ++k) {
        if (!live[k]) continue;
        if (zero) {
          part[k].y[sir_model::zero_first] = 0;
        }
        uint32_t err = 0;
        real ea = 0, eb = 0;
        sir_model::update<real>(dt_r, part[k].y, sh, rng[k], det, err, ea, eb);
        if (err) {
          live[k] = false;
          fail_mask |= 1u << k;
          sd.errors[p0 + k] =
              error_record{err, t, static_cast<double>(ea), static_cast<double>(eb)};
          atomicAdd(sd.error_count, 1u);
        }
      }
    }
#pragma unroll
    for (int k = 0; k < PPT; ++k) {
      w[k] = (p0 + k < P) ? sir_model::compare<real>(part[k].y, obs, lfact, sh, rng[k], det)
                          : 0.0;
    }
    time = time + n_steps * dt;

    int idx[PPT];
    const resample_result res = block_resample<BLOCK, PPT>(
        w, idx, P, cum, scratch, [&]() { return random_real<double>(frng); });
    if (tid == 0) {
      ll += res.ll_increment;
      fallbacks += res.n_fallback;
    }

    if (res.resampled) {
      // ancestor gather through LDS, one state component at a time
      real* xb = reinterpret_cast<real*>(cum);
#pragma unroll
      for (int c = 0; c < 5; ++c) {
        __syncthreads();
#pragma unroll
        for (int k = 0; k < PPT; ++k) {
          if (p0 + k < P) xb[p0 + k] = part[k].y[c];
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < PPT; ++k) {
          if (p0 + k < P) part[k].y[c] = xb[idx[k]];
        }
      }
      __syncthreads();
    }
    if (sd.ancestry) {
#pragma unroll
      for (int k = 0; k < PPT; ++k) {
        if (p0 + k < P) sd.ancestry[static_cast<size_t>(t) * P + p0 + k] = idx[k];
      }
    }
    if (flags & SLOT_SAVE_TRAJ) {
      real* tr = static_cast<real*>(sd.traj) + static_cast<size_t>(t) * 5 * P;
#pragma unroll
      for (int k = 0; k < PPT; ++k) {
        if (p0 + k < P) {
#pragma unroll
          for (int c = 0; c < 5; ++c) tr[c * P + p0 + k] = part[k].y[c];
        }
      }
    }
  }

#pragma unroll
  for (int k = 0; k < PPT; ++k) {
    const int p = p0 + k;
    if (p < P) {
#pragma unroll
      for (int c = 0; c < 5; ++c) state[c * P + p] = part[k].y[c];
      sd.rng[0 * rs + 2 + p] = rng[k].s0;
      sd.rng[1 * rs + 2 + p] = rng[k].s1;
      sd.rng[2 * rs + 2 + p] = rng[k].s2;
      sd.rng[3 * rs + 2 + p] = rng[k].s3;
    }
  }
  if (tid == 0) {
    sd.rng[0 * rs] = frng.s0;
    sd.rng[1 * rs] = frng.s1;
    sd.rng[2 * rs] = frng.s2;
    sd.rng[3 * rs] = frng.s3;
    sd.ll[0] = ll;
    sd.time[0] = time;
    if (fallbacks) atomicAdd(sd.n_fallback, fallbacks);
  }
}

template <typename real, int BLOCK, int PPT>
hipError_t launch_one(const filter_slot* d_slots, int n_slots, int max_particles,
                      hipStream_t stream) {
  const size_t lds = (kScratchDoubles + static_cast<size_t>(max_particles)) * sizeof(double);
  hipLaunchKernelGGL((sir_filter_kernel<real, BLOCK, PPT>), dim3(n_slots), dim3(BLOCK), lds,
                     stream, d_slots);
  return hipGetLastError();
}

template <typename real>
hipError_t dispatch(const filter_slot* d_slots, int n_slots, int max_particles,
                    launch_config cfg, hipStream_t stream) {
#define D2_CASE(B, K)                                                     \
  if (cfg.block == B && cfg.ppt == K) {                                   \
    return launch_one<real, B, K>(d_slots, n_slots, max_particles, stream); \
  }
  D2_CASE(64, 1)
  D2_CASE(64, 2)
  D2_CASE(64, 4)
  D2_CASE(128, 2)
  D2_CASE(128, 4)
  D2_CASE(256, 2)
  D2_CASE(256, 4)
  D2_CASE(256, 8)
  D2_CASE(512, 2)
  D2_CASE(512, 4)
  D2_CASE(512, 8)
  D2_CASE(1024, 2)
  D2_CASE(1024, 4)
  D2_CASE(1024, 8)
#undef D2_CASE
  return hipErrorInvalidConfiguration;
}

}  // namespace

bool sir_filter_config_valid(launch_config cfg, bool) {
  static const int ok[][2] = {{64, 1},   {64, 2},   {64, 4},   {128, 2}, {128, 4},
                              {256, 2},  {256, 4},  {256, 8},  {512, 2}, {512, 4},
                              {512, 8},  {1024, 2}, {1024, 4}, {1024, 8}};
  for (const auto& c : ok) {
    if (c[0] == cfg.block && c[1] == cfg.ppt) return true;
  }
  return false;
}

int sir_filter_max_particles(bool) { return 1024 * 8; }

launch_config sir_filter_default_config(int n_particles, bool) {
  if (n_particles <= 64) return {64, 1};
  if (n_particles <= 128) return {64, 2};
  if (n_particles <= 256) return {64, 4};
  if (n_particles <= 512) return {128, 4};
  if (n_particles <= 1024) return {512, 2};
  if (n_particles <= 2048) return {1024, 2};
  if (n_particles <= 4096) return {1024, 4};
  return {1024, 8};
}

hipError_t launch_sir_filter(const filter_slot* d_slots, int n_slots, int max_particles,
                             bool fp64, launch_config cfg, hipStream_t stream) {
  if (n_slots == 0) return hipSuccess;
  if (cfg.block * cfg.ppt < max_particles) return hipErrorInvalidConfiguration;
  return fp64 ? dispatch<double>(d_slots, n_slots, max_particles, cfg, stream)
              : dispatch<float>(d_slots, n_slots, max_particles, cfg, stream);
}

}  // namespace dust2_amd
