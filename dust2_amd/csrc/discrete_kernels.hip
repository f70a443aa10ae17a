// Discrete-time kernels templated on the model (models.hpp), one engine
// for any model of the contract, as dust_discrete<T> is
// (dust2/discrete/system.hpp:20-483):
//
//   advance_kernel     interval pipeline: gather the particle's state from
//                      its ancestor slot (the reorder of
//                      dust2/discrete/system.hpp:169-181 fused into the
//                      load), run the interval's steps (run_particle,
//                      :404-423, with zero_every), draw compare_data
//                      (:252-287) and write the FP64 log-weight;
//   filter_kernel      persistent filter: one workgroup owns a group for a
//                      whole run, state and streams in registers
//                      (dust2/filter.hpp:42-104 in one launch);
//   system_run_kernel, system_compare_kernel
//                      run_to_time (:61-92) and compare_data of a system
//                      whose state stays in HBM between calls.
//
// Models whose particles share a step stage across a wavefront (coupled:
// SEIR16's force of infection on the matrix cores) run in one-wavefront
// workgroups and keep every lane in the step loop; the others run one
// particle per thread in 256-thread workgroups and leave the loop at their
// first error.
//
// Layouts: pipeline and filter state real [S][P] per group, streams
// [4][P + 2] per group (2 filter streams, then the particles); system
// state [G][S][P], streams [G][4][P].
#include <hip/hip_runtime.h>

#include "block_resample.hpp"
#include "device_random.hpp"
#include "engine.hpp"
#include "interpolate.hpp"
#include "models.hpp"

namespace dust2_amd {

namespace {

template <class Model, typename real>
struct launch_shape {
  static constexpr bool coupled = Model::template coupled<real>;
  static constexpr int block = coupled ? 64 : 256;
};

template <class Model>
__device__ __forceinline__ void zero_every(double* y) {
#pragma unroll
  for (int z = 0; z < Model::zero_count; ++z) y[Model::zero_first + z] = 0;
}
template <class Model>
__device__ __forceinline__ void zero_every(float* y) {
#pragma unroll
  for (int z = 0; z < Model::zero_count; ++z) y[Model::zero_first + z] = 0;
}

// ---- interval pipeline -------------------------------------------------------
#ifndef DUST2_COUPLED_WAVES_PER_SIMD
#define DUST2_COUPLED_WAVES_PER_SIMD 3
#endif

// Workgroups map onto (slot, block of the slot's particles) either as the
// 2-D grid (block, slot), or, with a block map, as entry blockIdx.x of the
// map (slot << kBlockBits | block; kNoBlock: nothing to do), so that slots
// of very different particle counts share one launch without idle
// workgroups.
template <class Model, typename real, int BLOCK, int MINW>
__global__ void __launch_bounds__(BLOCK, MINW)
    advance_kernel(const pipe_slot* __restrict__ slots, const uint32_t* __restrict__ block_map,
                   int t, int parity, uint32_t mode) {
  constexpr bool coupled = Model::template coupled<real>;
  __shared__ float lds[coupled ? Model::lds_floats : 1];
  int slot = blockIdx.y;
  int blk = blockIdx.x;
  if (block_map) {
    const uint32_t e = block_map[blockIdx.x];
    if (e == kNoBlock) return;
    slot = static_cast<int>(e >> kBlockBits);
    blk = static_cast<int>(e & ((1u << kBlockBits) - 1));
  }
  const pipe_slot& sd = slots[slot];
  const int P = static_cast<int>(sd.n_particles);
  const int p = blk * BLOCK + threadIdx.x;
  const int lane = threadIdx.x % 64;
  if (coupled) {
    if (blk * BLOCK >= P) return;  // whole wavefront idle
  } else if (p >= P) {
    return;
  }
  if (!pipe_slot_interval(sd, t, parity, mode)) return;
  // the reference stops at the first interval with errors
  // (dust2/discrete/system.hpp:87): later intervals of the run do nothing
  if ((mode & PIPE_STEP) && __atomic_load_n(sd.error_count, __ATOMIC_RELAXED) != 0) return;
  const bool active = p < P;
  const bool det = (sd.flags & SLOT_DETERMINISTIC) != 0;
  const auto sh = Model::template load<real>(sd.pars);
  const int S = Model::size(sh);
  const real* state_in = static_cast<const real*>(parity ? sd.state_b : sd.state_a);
  real* state_out = static_cast<real*>(parity ? sd.state_a : sd.state_b);
  const int rs = P + 2;
  const bool draws = (mode & (PIPE_STEP | PIPE_INITIAL)) != 0;

  rng256 r{0, 0, 0, 0};
  if (active && draws) {
    r = rng256{sd.rng[2 + p], sd.rng[rs + 2 + p], sd.rng[2 * rs + 2 + p],
               sd.rng[3 * rs + 2 + p]};
  }
  real y[Model::n_state];
  if (mode & PIPE_INITIAL) {
#pragma unroll
    for (int c = 0; c < Model::n_state; ++c) y[c] = 0;
    if (active) Model::initial(sh, r, det, y);
  } else {
    const int src = active ? ((mode & PIPE_GATHER) ? static_cast<int>(sd.ancestors[p]) : p) : 0;
#pragma unroll
    for (int c = 0; c < Model::n_state; ++c) {
      y[c] = (active && c < S) ? state_in[static_cast<size_t>(c) * P + src] : real(0);
    }
  }
  if ((mode & PIPE_TRAJ_PREV) && sd.traj && active) {
    // post-reorder state of the previous interval (trajectories_.add)
    real* tr = static_cast<real*>(sd.traj) + static_cast<size_t>(t - 1) * S * P + p;
#pragma unroll
    for (int c = 0; c < Model::n_state; ++c) {
      if (c < S) tr[static_cast<size_t>(c) * P] = y[c];
    }
  }
  if (mode & PIPE_STEP) {
    const int n_steps = static_cast<int>(sd.sched_steps[t]);
    const uint64_t zero_mask = sd.sched_zero[t];
    const real dt_r = static_cast<real>(sd.dt);
    bool failed = !active;
    for (int s = 0; s < n_steps; ++s) {
      if (!coupled && failed) break;
      if (Model::zero_count > 0) {
        // steps beyond 64 are zeroed by the host's rule replicated here
        const bool zero = s < 64 ? ((zero_mask >> s) & 1u) != 0
                                 : fmod(sd.sched_time[t] + s * sd.dt, Model::zero_period) == 0;
        if (zero) zero_every<Model>(y);
      }
      auto sh_s = sh;
      if constexpr (Model::has_beta_schedule) {
        if (sd.beta_n > 0 && !failed) {
          // beta at the step's start time, as the step's update sees it
          const double time_s = sd.sched_time[t] + s * sd.dt;
          int ie = INTERP_OK;
          const double b = interp_eval(sd.beta_kind, sd.beta_t, sd.beta_y, sd.beta_k,
                                       sd.beta_n, 1, 0, time_s, ie);
          if (ie != INTERP_OK) {
            const double knot = ie == INTERP_BEFORE ? sd.beta_t[0] : sd.beta_t[sd.beta_n - 1];
            sd.errors[p] = error_record{ie == INTERP_BEFORE ? ERR_INTERP_BEFORE
                                                            : ERR_INTERP_AFTER,
                                        static_cast<uint32_t>(t), time_s, knot};
            atomicAdd(sd.error_count, 1u);
            failed = true;
            continue;
          }
          sh_s.beta = static_cast<real>(b);
        }
      }
      if constexpr (coupled) Model::prepare(y, sh_s, lds, lane);
      if (!failed) {
        uint32_t err = 0;
        real ea = 0, eb = 0;
        Model::update(dt_r, y, sh_s, lds, lane, r, det, err, ea, eb);
        if (err) {
          sd.errors[p] = error_record{err, static_cast<uint32_t>(t), static_cast<double>(ea),
                                      static_cast<double>(eb)};
          atomicAdd(sd.error_count, 1u);
          failed = true;
        }
      }
    }
    if (active) {
      // compare_data: the FP64 log-weight of the interval (NaN data draw
      // nothing and score 0)
      const size_t o = static_cast<size_t>(t) * sd.obs_stride;
      const double w = failed ? -__builtin_inf()
                              : Model::compare(y, sd.obs + o, sd.lfact + o, sh, r, det);
      sd.lambda[p] = w;
    }
  }
  if (active) {
    if (draws) {
      sd.rng[2 + p] = r.s0;
      sd.rng[rs + 2 + p] = r.s1;
      sd.rng[2 * rs + 2 + p] = r.s2;
      sd.rng[3 * rs + 2 + p] = r.s3;
    }
#pragma unroll
    for (int c = 0; c < Model::n_state; ++c) {
      if (c < S) state_out[static_cast<size_t>(c) * P + p] = y[c];
    }
  }
}

template <class Model, typename real>
hipError_t launch_advance_t(const pipe_slot* d_slots, int n_slots, int max_particles, int t,
                            int parity, uint32_t mode, hipStream_t stream) {
  constexpr int B = launch_shape<Model, real>::block;
  constexpr int W = launch_shape<Model, real>::coupled ? DUST2_COUPLED_WAVES_PER_SIMD : 1;
  const dim3 grid((max_particles + B - 1) / B, n_slots);
  hipLaunchKernelGGL((advance_kernel<Model, real, B, W>), grid, dim3(B), 0, stream, d_slots,
                     nullptr, t, parity, mode);
  return hipGetLastError();
}

template <class Model, typename real>
hipError_t launch_advance_mapped_t(const pipe_slot* d_slots, const uint32_t* d_block_map,
                                   int n_blocks, int t, hipStream_t stream) {
  constexpr int B = launch_shape<Model, real>::block;
  constexpr int W = launch_shape<Model, real>::coupled ? DUST2_COUPLED_WAVES_PER_SIMD : 1;
  hipLaunchKernelGGL((advance_kernel<Model, real, B, W>), dim3(n_blocks), dim3(B), 0, stream,
                     d_slots, d_block_map, t, 0, 0u);
  return hipGetLastError();
}

// ---- persistent filter ---------------------------------------------------------
// Thread t owns particles t*PPT .. t*PPT+PPT-1 (their slots, and the slots'
// streams: streams stay with slots on reorder, vignettes/design.Rmd:133).
// The thread's particles advance step by step together, so the
// independent sampler chains of its PPT particles interleave.  For
// uncoupled models whose state fits in registers.
template <class Model, typename real, int BLOCK, int PPT>
__global__ void __launch_bounds__(BLOCK) filter_kernel(const filter_slot* __restrict__ slots) {
  static_assert(!Model::template coupled<real>, "the persistent filter runs uncoupled models");
  constexpr int NS = Model::n_state;
  extern __shared__ double smem[];
  double* scratch = smem;
  double* cum = smem + kScratchDoubles;

  const filter_slot& sd = slots[blockIdx.x];
  const int P = static_cast<int>(sd.n_particles);
  const int tid = threadIdx.x;
  const int p0 = tid * PPT;
  const uint32_t flags = sd.flags;
  const bool det = (flags & SLOT_DETERMINISTIC) != 0;
  const auto sh = Model::template load<real>(sd.pars);
  const int S = Model::size(sh);
  const double dt = sd.dt;
  const real dt_r = static_cast<real>(dt);
  real* state = static_cast<real*>(sd.state);
  const int rs = P + 2;  // rng stride: 2 filter streams then P particles

  real y[PPT][NS];
  rng256 rng[PPT];
  uint32_t fail_mask = 0;
#pragma unroll
  for (int k = 0; k < PPT; ++k) {
    const int p = p0 + k;
#pragma unroll
    for (int c = 0; c < NS; ++c) y[k][c] = 0;
    rng[k] = rng256{0, 0, 0, 0};
    if (p < P) {
      rng[k].s0 = sd.rng[0 * rs + 2 + p];
      rng[k].s1 = sd.rng[1 * rs + 2 + p];
      rng[k].s2 = sd.rng[2 * rs + 2 + p];
      rng[k].s3 = sd.rng[3 * rs + 2 + p];
      if (flags & SLOT_SET_INITIAL) {
        Model::initial(sh, rng[k], det, y[k]);
      } else {
#pragma unroll
        for (int c = 0; c < NS; ++c) {
          if (c < S) y[k][c] = state[c * P + p];
        }
      }
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
    if (Model::zero_count > 0 && !zero_generic) {
      for (int s = 0; s < n_steps; ++s) {
        if (fmod(time + s * dt, Model::zero_period) == 0) zero_mask |= uint64_t(1) << s;
      }
    }

    double w[PPT];
    bool live[PPT];
#pragma unroll
    for (int k = 0; k < PPT; ++k) live[k] = (p0 + k < P) && !((fail_mask >> k) & 1u);
    for (int s = 0; s < n_steps; ++s) {
      const bool zero = Model::zero_count > 0 &&
                        (zero_generic ? fmod(time + s * dt, Model::zero_period) == 0
                                      : ((zero_mask >> s) & 1u) != 0);
#pragma unroll
      for (int k = 0; k < PPT; ++k) {
        if (!live[k]) continue;
        if (zero) zero_every<Model>(y[k]);
        uint32_t err = 0;
        real ea = 0, eb = 0;
        Model::update(dt_r, y[k], sh, nullptr, 0, rng[k], det, err, ea, eb);
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
      w[k] = (p0 + k < P) ? Model::compare(y[k], obs, lfact, sh, rng[k], det) : 0.0;
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
      for (int c = 0; c < NS; ++c) {
        if (c >= S) break;
        __syncthreads();
#pragma unroll
        for (int k = 0; k < PPT; ++k) {
          if (p0 + k < P) xb[p0 + k] = y[k][c];
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < PPT; ++k) {
          if (p0 + k < P) y[k][c] = xb[idx[k]];
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
      real* tr = static_cast<real*>(sd.traj) + static_cast<size_t>(t) * S * P;
#pragma unroll
      for (int k = 0; k < PPT; ++k) {
        if (p0 + k < P) {
#pragma unroll
          for (int c = 0; c < NS; ++c) {
            if (c < S) tr[c * P + p0 + k] = y[k][c];
          }
        }
      }
    }
  }

#pragma unroll
  for (int k = 0; k < PPT; ++k) {
    const int p = p0 + k;
    if (p < P) {
#pragma unroll
      for (int c = 0; c < NS; ++c) {
        if (c < S) state[c * P + p] = y[k][c];
      }
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

template <class Model, typename real, int BLOCK, int PPT>
hipError_t launch_filter_one(const filter_slot* d_slots, int n_slots, int max_particles,
                             hipStream_t stream) {
  const size_t lds = (kScratchDoubles + static_cast<size_t>(max_particles)) * sizeof(double);
  hipLaunchKernelGGL((filter_kernel<Model, real, BLOCK, PPT>), dim3(n_slots), dim3(BLOCK), lds,
                     stream, d_slots);
  return hipGetLastError();
}

// The persistent filter's launch shapes (threads x particles per thread).
constexpr int kFilterShapes[][2] = {{64, 1},   {64, 2},   {64, 4},   {128, 2}, {128, 4},
                                    {256, 2},  {256, 4},  {256, 8},  {512, 2}, {512, 4},
                                    {512, 8},  {1024, 2}, {1024, 4}, {1024, 8}};

template <class Model, typename real>
hipError_t dispatch_filter(const filter_slot* d_slots, int n_slots, int max_particles,
                           launch_config cfg, hipStream_t stream) {
#define D2_CASE(B, K)                                                                     \
  if (cfg.block == B && cfg.ppt == K) {                                                   \
    return launch_filter_one<Model, real, B, K>(d_slots, n_slots, max_particles, stream); \
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

// ---- system ----------------------------------------------------------------------
template <class Model, typename real, int BLOCK, int MINW>
__global__ void __launch_bounds__(BLOCK, MINW)
    system_run_kernel(real* __restrict__ state, uint64_t* __restrict__ rng,
                      const double* __restrict__ kpars, int kstride, int G, int P, double time,
                      double dt, int n_steps, uint32_t flags, error_record* errors,
                      uint32_t* error_count) {
  constexpr bool coupled = Model::template coupled<real>;
  __shared__ float lds[coupled ? Model::lds_floats : 1];
  const int g = blockIdx.y;
  const int p = blockIdx.x * BLOCK + threadIdx.x;
  const int lane = threadIdx.x % 64;
  if (g >= G) return;
  if (coupled) {
    if (static_cast<int>(blockIdx.x) * BLOCK >= P) return;
  } else if (p >= P) {
    return;
  }
  const bool active = p < P;
  const auto sh = Model::template load<real>(kpars + static_cast<size_t>(kstride) * g);
  const int S = Model::size(sh);
  real* st = state + static_cast<size_t>(g) * S * P;
  uint64_t* rg = rng + static_cast<size_t>(g) * 4 * P;
  const bool det = (flags & SLOT_DETERMINISTIC) != 0;
  rng256 r{0, 0, 0, 0};
  if (active) r = rng256{rg[p], rg[P + p], rg[2 * P + p], rg[3 * P + p]};
  real y[Model::n_state];
#pragma unroll
  for (int c = 0; c < Model::n_state; ++c) y[c] = 0;
  if (active) {
    if (flags & SLOT_SET_INITIAL) {
      Model::initial(sh, r, det, y);
    } else {
#pragma unroll
      for (int c = 0; c < Model::n_state; ++c) {
        if (c < S) y[c] = st[static_cast<size_t>(c) * P + p];
      }
    }
  }
  const real dt_r = static_cast<real>(dt);
  bool failed = !active;
  for (int s = 0; s < n_steps; ++s) {
    if (!coupled && failed) break;
    const double time_s = time + s * dt;
    if (Model::zero_count > 0 && fmod(time_s, Model::zero_period) == 0) zero_every<Model>(y);
    if constexpr (coupled) Model::prepare(y, sh, lds, lane);
    if (!failed) {
      uint32_t err = 0;
      real ea = 0, eb = 0;
      Model::update(dt_r, y, sh, lds, lane, r, det, err, ea, eb);
      if (err) {
        errors[static_cast<size_t>(g) * P + p] =
            error_record{err, 0u, static_cast<double>(ea), static_cast<double>(eb)};
        atomicAdd(error_count, 1u);
        failed = true;
      }
    }
  }
  if (active) {
#pragma unroll
    for (int c = 0; c < Model::n_state; ++c) {
      if (c < S) st[static_cast<size_t>(c) * P + p] = y[c];
    }
    rg[p] = r.s0;
    rg[P + p] = r.s1;
    rg[2 * P + p] = r.s2;
    rg[3 * P + p] = r.s3;
  }
}

template <class Model, typename real>
__global__ void __launch_bounds__(256)
    system_compare_kernel(const real* __restrict__ state, uint64_t* __restrict__ rng,
                          const double* __restrict__ kpars, int kstride, const double* obs,
                          const double* lfact, int obs_group_stride, int G, int P,
                          uint32_t flags, double* out) {
  const int g = blockIdx.y;
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P || g >= G) return;
  const auto sh = Model::template load<real>(kpars + static_cast<size_t>(kstride) * g);
  const int S = Model::size(sh);
  const real* st = state + static_cast<size_t>(g) * S * P;
  uint64_t* rg = rng + static_cast<size_t>(g) * 4 * P;
  real y[Model::n_state];
#pragma unroll
  for (int c = 0; c < Model::n_state; ++c) {
    y[c] = c < S ? st[static_cast<size_t>(c) * P + p] : real(0);
  }
  rng256 r{rg[p], rg[P + p], rg[2 * P + p], rg[3 * P + p]};
  const int d = g * obs_group_stride;
  out[static_cast<size_t>(g) * P + p] =
      Model::compare(y, obs + d, lfact + d, sh, r, (flags & SLOT_DETERMINISTIC) != 0);
  rg[p] = r.s0;
  rg[P + p] = r.s1;
  rg[2 * P + p] = r.s2;
  rg[3 * P + p] = r.s3;
}

template <class Model, typename real>
hipError_t launch_run_t(const discrete_run_args& a, hipStream_t stream) {
  constexpr int B = launch_shape<Model, real>::block;
  constexpr int W = launch_shape<Model, real>::coupled ? DUST2_COUPLED_WAVES_PER_SIMD : 1;
  const dim3 grid((a.P + B - 1) / B, a.G);
  hipLaunchKernelGGL((system_run_kernel<Model, real, B, W>), grid, dim3(B), 0, stream,
                     static_cast<real*>(a.state), a.rng, a.kpars, a.kstride, a.G, a.P, a.time,
                     a.dt, a.n_steps, a.flags, a.errors, a.error_count);
  return hipGetLastError();
}

template <class Model, typename real>
hipError_t launch_compare_t(const discrete_run_args& a, const double* obs, const double* lfact,
                            int obs_group_stride, double* out, hipStream_t stream) {
  const dim3 grid((a.P + 255) / 256, a.G);
  hipLaunchKernelGGL((system_compare_kernel<Model, real>), grid, dim3(256), 0, stream,
                     static_cast<const real*>(a.state), a.rng, a.kpars, a.kstride, obs, lfact,
                     obs_group_stride, a.G, a.P, a.flags, out);
  return hipGetLastError();
}

}  // namespace

// ---- launchers, dispatched on the model id ------------------------------------------
hipError_t launch_discrete_advance(int model, bool fp64, const pipe_slot* d_slots, int n_slots,
                                   int max_particles, int t, int parity, uint32_t mode,
                                   hipStream_t stream) {
  if (n_slots == 0) return hipSuccess;
  switch (model) {
    case MODEL_SIR:
      return fp64 ? launch_advance_t<sir_model, double>(d_slots, n_slots, max_particles, t,
                                                         parity, mode, stream)
                  : launch_advance_t<sir_model, float>(d_slots, n_slots, max_particles, t,
                                                        parity, mode, stream);
    case MODEL_SEIR16:
      return fp64 ? launch_advance_t<seir16_model, double>(d_slots, n_slots, max_particles, t,
                                                            parity, mode, stream)
                  : launch_advance_t<seir16_model, float>(d_slots, n_slots, max_particles, t,
                                                           parity, mode, stream);
    case MODEL_WALK:
      return fp64 ? launch_advance_t<walk_model, double>(d_slots, n_slots, max_particles, t,
                                                          parity, mode, stream)
                  : launch_advance_t<walk_model, float>(d_slots, n_slots, max_particles, t,
                                                         parity, mode, stream);
    default:
      return hipErrorInvalidValue;
  }
}

int discrete_advance_block(int model, bool fp64) {
  switch (model) {
    case MODEL_SIR:
      return fp64 ? launch_shape<sir_model, double>::block : launch_shape<sir_model, float>::block;
    case MODEL_SEIR16:
      return fp64 ? launch_shape<seir16_model, double>::block
                  : launch_shape<seir16_model, float>::block;
    case MODEL_WALK:
      return fp64 ? launch_shape<walk_model, double>::block
                  : launch_shape<walk_model, float>::block;
    default:
      return 0;
  }
}

hipError_t launch_discrete_advance_mapped(int model, bool fp64, const pipe_slot* d_slots,
                                          const uint32_t* d_block_map, int n_blocks, int t,
                                          hipStream_t stream) {
  if (n_blocks == 0) return hipSuccess;
  switch (model) {
    case MODEL_SIR:
      return fp64 ? launch_advance_mapped_t<sir_model, double>(d_slots, d_block_map, n_blocks,
                                                                t, stream)
                  : launch_advance_mapped_t<sir_model, float>(d_slots, d_block_map, n_blocks, t,
                                                               stream);
    case MODEL_SEIR16:
      return fp64 ? launch_advance_mapped_t<seir16_model, double>(d_slots, d_block_map,
                                                                   n_blocks, t, stream)
                  : launch_advance_mapped_t<seir16_model, float>(d_slots, d_block_map, n_blocks,
                                                                  t, stream);
    case MODEL_WALK:
      return fp64 ? launch_advance_mapped_t<walk_model, double>(d_slots, d_block_map, n_blocks,
                                                                 t, stream)
                  : launch_advance_mapped_t<walk_model, float>(d_slots, d_block_map, n_blocks,
                                                                t, stream);
    default:
      return hipErrorInvalidValue;
  }
}

bool discrete_filter_config_valid(launch_config cfg) {
  for (const auto& c : kFilterShapes) {
    if (c[0] == cfg.block && c[1] == cfg.ppt) return true;
  }
  return false;
}

int discrete_filter_max_particles(int model) {
  return model == MODEL_SIR ? 1024 * 8 : 0;
}

launch_config discrete_filter_default_config(int n_particles) {
  if (n_particles <= 64) return {64, 1};
  if (n_particles <= 128) return {64, 2};
  if (n_particles <= 256) return {64, 4};
  if (n_particles <= 512) return {128, 4};
  if (n_particles <= 1024) return {512, 2};
  if (n_particles <= 2048) return {1024, 2};
  if (n_particles <= 4096) return {1024, 4};
  return {1024, 8};
}

hipError_t launch_discrete_filter(int model, bool fp64, const filter_slot* d_slots, int n_slots,
                                  int max_particles, launch_config cfg, hipStream_t stream) {
  if (n_slots == 0) return hipSuccess;
  if (cfg.block * cfg.ppt < max_particles) return hipErrorInvalidConfiguration;
  if (model != MODEL_SIR) return hipErrorInvalidValue;
  return fp64 ? dispatch_filter<sir_model, double>(d_slots, n_slots, max_particles, cfg, stream)
              : dispatch_filter<sir_model, float>(d_slots, n_slots, max_particles, cfg, stream);
}

hipError_t launch_discrete_run(int model, bool fp64, const discrete_run_args& a,
                               hipStream_t stream) {
  if (a.G == 0 || a.P == 0) return hipSuccess;
  switch (model) {
    case MODEL_SIR:
      return fp64 ? launch_run_t<sir_model, double>(a, stream)
                  : launch_run_t<sir_model, float>(a, stream);
    case MODEL_SEIR16:
      return fp64 ? launch_run_t<seir16_model, double>(a, stream)
                  : launch_run_t<seir16_model, float>(a, stream);
    case MODEL_WALK:
      return fp64 ? launch_run_t<walk_model, double>(a, stream)
                  : launch_run_t<walk_model, float>(a, stream);
    default:
      return hipErrorInvalidValue;
  }
}

hipError_t launch_discrete_compare(int model, bool fp64, const discrete_run_args& a,
                                   const double* obs, const double* lfact, int obs_group_stride,
                                   double* out, hipStream_t stream) {
  if (a.G == 0 || a.P == 0) return hipSuccess;
  switch (model) {
    case MODEL_SIR:
      return fp64 ? launch_compare_t<sir_model, double>(a, obs, lfact, obs_group_stride, out,
                                                         stream)
                  : launch_compare_t<sir_model, float>(a, obs, lfact, obs_group_stride, out,
                                                        stream);
    case MODEL_SEIR16:
      return fp64 ? launch_compare_t<seir16_model, double>(a, obs, lfact, obs_group_stride,
                                                            out, stream)
                  : launch_compare_t<seir16_model, float>(a, obs, lfact, obs_group_stride, out,
                                                           stream);
    default:
      return hipErrorInvalidValue;
  }
}

}  // namespace dust2_amd
