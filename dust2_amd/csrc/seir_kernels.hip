// Age-structured stochastic SEIR (16 age groups x S, E, I, R + 16 weekly
// incidence counters = 80 states) for the interval pipeline, FP32.
//
// No counterpart model exists in the reference; it is written to the same
// generated-class contract as inst/examples/sir.cpp (initial / update /
// compare_data / zero_every) and pinned by the CPU oracle's seir16 model.
//
// Force of infection: lambda[a][p] = sum_b A[a][b] * I[b][p] with
// A[a][b] = beta * C[a][b] / N[b] folded on the host.  For the 64
// particles of a wavefront this is a 16 x 16 x 64 contraction, done on the
// matrix cores as 4 column tiles x 4 k-chunks of v_mfma_f32_16x16x4f32:
//   A operand: lane l holds A[l % 16][4*kk + l / 16]
//   B operand: lane l holds I[4*kk + l / 16][16*tile + l % 16]
//   result:    lane l, element i holds lambda[4*(l / 16) + i][16*tile + l % 16]
// Each lane keeps its particle's 80 states in registers (the age loop is
// unrolled, so every index is a compile-time constant); LDS holds only the
// wavefront's 16 x 64 infectious counts and forces of infection, which give
// the MFMA operands their transposed layouts.  8 KB of LDS per wavefront
// leaves the occupancy to the register budget.
#include <hip/hip_runtime.h>

#include "device_random.hpp"
#include "engine.hpp"

namespace dust2_amd {

namespace {

typedef float v4f __attribute__((ext_vector_type(4)));

constexpr int kAge = 16;
constexpr int kSeirState = 5 * kAge;
constexpr int kWaveP = 64;

// Kernel parameter block (doubles): see seir_kernel_pars() on the host.
constexpr int kParEI = 1, kParIR = 2, kParNoise = 3, kParI0 = 4, kParN = 5,
              kParA = 5 + kAge;

#ifndef DUST2_SEIR_WAVES_PER_SIMD
#define DUST2_SEIR_WAVES_PER_SIMD 3
#endif
// the register budget for DUST2_SEIR_WAVES_PER_SIMD resident wavefronts
__global__ void __launch_bounds__(kWaveP, DUST2_SEIR_WAVES_PER_SIMD)
    seir_advance_kernel(const pipe_slot* __restrict__ slots, int t, int parity, uint32_t mode) {
  __shared__ float xi[kAge * kWaveP];   // I[a][lane]
  __shared__ float foi[kAge * kWaveP];  // lambda[a][lane]
  const pipe_slot& sd = slots[blockIdx.y];
  const int P = static_cast<int>(sd.n_particles);
  const int lane = threadIdx.x;
  const int p = blockIdx.x * kWaveP + lane;
  if (static_cast<int>(blockIdx.x) * kWaveP >= P) return;  // whole wavefront idle
  if (!pipe_slot_interval(sd, t, parity, mode)) return;
  if ((mode & PIPE_STEP) && __atomic_load_n(sd.error_count, __ATOMIC_RELAXED) != 0) return;
  const bool active = p < P;
  const bool det = (sd.flags & SLOT_DETERMINISTIC) != 0;
  const double* pr = sd.pars;
  const float* state_in = static_cast<const float*>(parity ? sd.state_b : sd.state_a);
  float* state_out = static_cast<float*>(parity ? sd.state_a : sd.state_b);

  float S[kAge], E[kAge], I[kAge], R[kAge], inc[kAge];
  if (mode & PIPE_INITIAL) {
    const float I0 = static_cast<float>(pr[kParI0]);
#pragma unroll
    for (int a = 0; a < kAge; ++a) {
      const float N = static_cast<float>(pr[kParN + a]);
      S[a] = active ? N - I0 : 0.f;
      E[a] = 0.f;
      I[a] = active ? I0 : 0.f;
      R[a] = 0.f;
      inc[a] = 0.f;
    }
  } else {
    const int src = active ? ((mode & PIPE_GATHER) ? static_cast<int>(sd.ancestors[p]) : p) : 0;
    const float* x = state_in + src;
    const size_t Pz = static_cast<size_t>(P);
#pragma unroll
    for (int a = 0; a < kAge; ++a) {
      S[a] = active ? x[a * Pz] : 0.f;
      E[a] = active ? x[(kAge + a) * Pz] : 0.f;
      I[a] = active ? x[(2 * kAge + a) * Pz] : 0.f;
      R[a] = active ? x[(3 * kAge + a) * Pz] : 0.f;
      inc[a] = active ? x[(4 * kAge + a) * Pz] : 0.f;
    }
  }
  if ((mode & PIPE_TRAJ_PREV) && sd.traj && active) {
    float* tr = static_cast<float*>(sd.traj) + static_cast<size_t>(t - 1) * kSeirState * P + p;
    const size_t Pz = static_cast<size_t>(P);
#pragma unroll
    for (int a = 0; a < kAge; ++a) {
      tr[a * Pz] = S[a];
      tr[(kAge + a) * Pz] = E[a];
      tr[(2 * kAge + a) * Pz] = I[a];
      tr[(3 * kAge + a) * Pz] = R[a];
      tr[(4 * kAge + a) * Pz] = inc[a];
    }
  }

  if (mode & PIPE_STEP) {
    const int rs = P + 2;
    rng256 r{0, 0, 0, 0};
    if (active) {
      r = rng256{sd.rng[2 + p], sd.rng[rs + 2 + p], sd.rng[2 * rs + 2 + p],
                 sd.rng[3 * rs + 2 + p]};
    }
    // A operand of the four k-chunks, fixed for the group
    float a_op[4];
#pragma unroll
    for (int kk = 0; kk < 4; ++kk) {
      a_op[kk] = static_cast<float>(pr[kParA + (lane % 16) * kAge + 4 * kk + lane / 16]);
    }
    const float p_EI = static_cast<float>(pr[kParEI]);
    const float p_IR = static_cast<float>(pr[kParIR]);
    const float dt = static_cast<float>(sd.dt);
    const int n_steps = static_cast<int>(sd.sched_steps[t]);
    const uint64_t zero_mask = sd.sched_zero[t];
    bool failed = false;
    for (int s = 0; s < n_steps; ++s) {
      const bool zero = s < 64 ? ((zero_mask >> s) & 1u) != 0
                               : fmod(sd.sched_time[t] + s * sd.dt, 7.0) == 0;
      if (zero) {
#pragma unroll
        for (int a = 0; a < kAge; ++a) inc[a] = 0.f;
      }
#pragma unroll
      for (int a = 0; a < kAge; ++a) xi[a * kWaveP + lane] = I[a];
      __syncthreads();
      // force of infection on the matrix cores (all 64 lanes take part)
#pragma unroll
      for (int tile = 0; tile < 4; ++tile) {
        v4f d = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) {
          const float b = xi[(4 * kk + lane / 16) * kWaveP + tile * 16 + lane % 16];
          d = __builtin_amdgcn_mfma_f32_16x16x4f32(a_op[kk], b, d, 0, 0, 0);
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          foi[(4 * (lane / 16) + i) * kWaveP + tile * 16 + lane % 16] = d[i];
        }
      }
      __syncthreads();
      if (active && !failed) {
#pragma unroll
        for (int a = 0; a < kAge; ++a) {
          if (failed) continue;
          const float lam = foi[a * kWaveP + lane];
          const float p_SE = prob_from_rate(-lam * dt);
          uint32_t err = 0;
          float bad_n = 0, bad_p = 0;
          const float n_SE = binomial<float>(r, S[a], p_SE, det, err);
          if (err) { bad_n = S[a]; bad_p = p_SE; }
          const float n_EI = err ? 0.f : binomial<float>(r, E[a], p_EI, det, err);
          if (err && bad_n == 0 && bad_p == 0) { bad_n = E[a]; bad_p = p_EI; }
          const float n_IR = err ? 0.f : binomial<float>(r, I[a], p_IR, det, err);
          if (err && bad_n == 0 && bad_p == 0) { bad_n = I[a]; bad_p = p_IR; }
          if (err) {
            sd.errors[p] = error_record{err, static_cast<uint32_t>(t), bad_n, bad_p};
            atomicAdd(sd.error_count, 1u);
            failed = true;
            continue;
          }
          S[a] = S[a] - n_SE;
          E[a] = E[a] + n_SE - n_EI;
          I[a] = I[a] + n_EI - n_IR;
          R[a] += n_IR;
          inc[a] += n_SE;
        }
      }
    }
    // compare: sum over age groups of log Poisson(obs; inc + Exp(noise)),
    // missing ages skipped without a draw; the FP64 log-weight is written
    // directly (SLOT_WEIGHT_IS_LL)
    if (active) {
      const double* obs = sd.obs + static_cast<size_t>(t) * sd.obs_stride;
      const double* lf = sd.lfact + static_cast<size_t>(t) * sd.obs_stride;
      const float rate = static_cast<float>(pr[kParNoise]);
      double w = 0;
#pragma unroll
      for (int a = 0; a < kAge; ++a) {
        const double x = obs[a];
        if (isnan(x)) continue;
        const double noise = static_cast<double>(exponential_rate<float>(r, rate, det));
        const double lambda = static_cast<double>(inc[a]) + noise;
        w += dpois_log(x, lambda, lf[a]);
      }
      sd.lambda[p] = failed ? -__builtin_inf() : w;
      sd.rng[2 + p] = r.s0;
      sd.rng[rs + 2 + p] = r.s1;
      sd.rng[2 * rs + 2 + p] = r.s2;
      sd.rng[3 * rs + 2 + p] = r.s3;
    }
  }
  if (active) {
    float* y = state_out + p;
    const size_t Pz = static_cast<size_t>(P);
#pragma unroll
    for (int a = 0; a < kAge; ++a) {
      y[a * Pz] = S[a];
      y[(kAge + a) * Pz] = E[a];
      y[(2 * kAge + a) * Pz] = I[a];
      y[(3 * kAge + a) * Pz] = R[a];
      y[(4 * kAge + a) * Pz] = inc[a];
    }
  }
}

}  // namespace

hipError_t launch_seir_advance(const pipe_slot* d_slots, int n_slots, int max_particles, int t,
                               int parity, uint32_t mode, hipStream_t stream) {
  if (n_slots == 0) return hipSuccess;
  const dim3 grid((max_particles + kWaveP - 1) / kWaveP, n_slots);
  hipLaunchKernelGGL(seir_advance_kernel, grid, dim3(kWaveP), 0, stream, d_slots, t, parity,
                     mode);
  return hipGetLastError();
}

}  // namespace dust2_amd
