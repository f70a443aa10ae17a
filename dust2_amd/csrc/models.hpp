// Models compiled as device functions, written to dust2's generated-class
// contract: initial / update / compare_data / rhs with shared (per group,
// read-only) parameters and a zero_every reset rule.  An odin2-generated
// class maps onto this shape with its build_shared/build_data kept on the
// host (dust2/properties.hpp:13-107 detects the same members).
//
// Discrete-time models are run by one set of kernels templated on the
// model (discrete_kernels.hip: the interval pipeline, the persistent filter
// and the system), as dust_discrete<T> is one engine for any T
// (dust2/discrete/system.hpp:20-483).  A discrete-time model provides:
//
//   n_state              states per particle (an upper bound when the
//                        length is a parameter; size() gives the actual one)
//   n_data               data values per time (0: no compare_data)
//   zero_period, zero_first, zero_count   the zero_every rule
//                        (dust2/properties.hpp:161-176); zero_count 0: none
//   shared<real>         the per-group shared state, built on the device by
//   load<real>(kp)       from the host-made kernel parameter block kp
//   size(sh)             states of this group
//   initial(sh, rng, det, y)
//   update(dt, y, sh, lds, lane, rng, det, err, err_a, err_b)
//                        one step in place; on invalid input sets err and
//                        the two values the error message reports
//   compare(y, obs, lfact, sh, rng, det)
//                        the FP64 log-weight of one particle; NaN data
//                        score 0 and draw nothing
//   coupled<real>        true when the particles of a wavefront compute a
//                        stage of every step together (prepare, through
//                        LDS, lds_floats per wavefront); their kernels keep
//                        every lane of the wavefront in the step loop
//   has_beta_schedule    beta may follow an interpolated schedule in time
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device_random.hpp"

namespace dust2_amd {

// Model ids shared by the host runtime and the kernel launchers.
enum discrete_model_id : int { MODEL_SIR = 0, MODEL_SEIR16 = 1, MODEL_WALK = 2 };

// Stochastic SIR, inst/examples/sir.cpp:12-199.
// Host parameter row: N, I0, beta, gamma, exp_noise.  Kernel block (8
// doubles): N, I0, beta, gamma, exp_noise, p_IR (FP64), p_IR (FP32), -.
struct sir_model {
  static constexpr int n_state = 5;
  static constexpr int n_pars = 5;
  static constexpr int n_data = 1;
  // zero_every = {{1, {4}}}: daily incidence reset (sir.cpp:95-97)
  static constexpr double zero_period = 1;
  static constexpr int zero_first = 4;
  static constexpr int zero_count = 1;
  static constexpr bool has_beta_schedule = true;
  static constexpr int lds_floats = 0;
  template <typename real>
  static constexpr bool coupled = false;

  template <typename real>
  struct shared {
    real N, I0, beta, gamma;
    real p_IR;         // 1 - exp(-gamma*dt), computed on the host
    double exp_noise;  // rate of the compare noise
  };

  template <typename real>
  __device__ __forceinline__ static shared<real> load(const double* kp) {
    shared<real> sh;
    sh.N = static_cast<real>(kp[0]);
    sh.I0 = static_cast<real>(kp[1]);
    sh.beta = static_cast<real>(kp[2]);
    sh.gamma = static_cast<real>(kp[3]);
    sh.exp_noise = kp[4];
    sh.p_IR = static_cast<real>(sizeof(real) == 8 ? kp[5] : kp[6]);
    return sh;
  }

  template <typename real>
  __device__ __forceinline__ static int size(const shared<real>&) {
    return n_state;
  }

  template <typename real>
  __device__ __forceinline__ static void initial(const shared<real>& sh, rng256&, bool,
                                                 real* y) {
    y[0] = sh.N - sh.I0;
    y[1] = sh.I0;
    y[2] = 0;
    y[3] = 0;
    y[4] = 0;
  }

  // -beta * I / N * dt.  FP64 in the reference's association,
  // ((-beta * I) / N) * dt; FP32 as -I times the step-invariant
  // beta * dt / N, with no division per step.
  __device__ __forceinline__ static double rate_SI(const shared<double>& sh, double I,
                                                   double dt) {
    return -sh.beta * I / sh.N * dt;
  }
  __device__ __forceinline__ static float rate_SI(const shared<float>& sh, float I, float dt) {
    return -I * (sh.beta * (dt * __builtin_amdgcn_rcpf(sh.N)));
  }

  template <typename real>
  __device__ __forceinline__ static void update(real dt, real* y, const shared<real>& sh,
                                                const float*, int, rng256& rng, bool det,
                                                uint32_t& err, real& err_a, real& err_b) {
    const real S = y[0];
    const real I = y[1];
    const real p_SI = prob_from_rate(rate_SI(sh, I, dt));
    const real n_SI = binomial<real>(rng, S, p_SI, det, err);
    if (err) {
      err_a = S;
      err_b = p_SI;
      return;
    }
    const real n_IR = binomial<real>(rng, I, sh.p_IR, det, err);
    if (err) {
      err_a = I;
      err_b = sh.p_IR;
      return;
    }
    y[0] = S - n_SI;
    y[1] = I + n_SI - n_IR;
    y[2] = y[2] + n_IR;
    y[3] = y[3] + n_SI;
    y[4] = y[4] + n_SI;
  }

  // Poisson(incidence + Exp(exp_noise)) log-density in FP64; NaN data
  // draws nothing and scores 0 (sir.cpp:105-120).
  template <typename real>
  __device__ __forceinline__ static double compare(const real* y, const double* obs,
                                                   const double* lfact, const shared<real>& sh,
                                                   rng256& rng, bool det) {
    const double x = obs[0];
    if (isnan(x)) {
      return 0;
    }
    double noise;
    if (sizeof(real) == 8) {
      noise = exponential_rate<double>(rng, sh.exp_noise, det);
    } else {
      noise = static_cast<double>(
          exponential_rate<float>(rng, static_cast<float>(sh.exp_noise), det));
    }
    const double lambda = static_cast<double>(y[4]) + noise;
    return dpois_log(x, lambda, lfact[0]);
  }
};

// Age-structured stochastic SEIR: 16 age groups x (S, E, I, R) + 16
// incidence counters reset every 7 days (weekly observations per age
// group).  No counterpart model exists in the reference; it is written to
// the same contract as sir.  Force of infection
// lambda_a = beta * sum_b C[a][b] I_b / N_b; binomial S->E, E->I, I->R per
// age; compare: sum over ages of log Poisson(obs_a; inc_a + Exp(exp_noise)).
//
// FP64 follows the CPU arithmetic order term by term (x_b = I_b / N_b,
// then the sum over b in order, then 1 - exp(-beta * foi * dt)), so draws
// reproduce.  FP32 computes the force of infection of the 64 particles of
// a wavefront on the matrix cores (coupled): lambda[a][p] =
// sum_b A[a][b] I[b][p] with A[a][b] = beta C[a][b] / N[b] folded on the
// host, a 16 x 16 x 64 contraction done as 4 column tiles x 4 k-chunks of
// v_mfma_f32_16x16x4f32:
//   A operand: lane l holds A[l % 16][4*kk + l / 16]
//   B operand: lane l holds I[4*kk + l / 16][16*tile + l % 16]
//   result:    lane l, element i holds lambda[4*(l / 16) + i][16*tile + l % 16]
// LDS holds the wavefront's 16 x 64 infectious counts and forces of
// infection (8 KB), which gives the operands their transposed layouts.
//
// Kernel block (doubles): [0] beta, [1] p_EI and [2] p_IR as FP32 values
// of 1 - exp(-rate*dt), [3] exp_noise, [4] I0, [5..20] N,
// [21 + 16a + b] beta C[a][b] / N[b], [277] p_EI and [278] p_IR in FP64
// (1 - exp(-rate*dt)), [279 + 16a + b] C[a][b].
struct seir16_model {
  static constexpr int n_age = 16;
  static constexpr int n_state = 5 * n_age;
  static constexpr int n_data = n_age;
  static constexpr double zero_period = 7;
  static constexpr int zero_first = 4 * n_age;
  static constexpr int zero_count = n_age;
  static constexpr bool has_beta_schedule = false;
  static constexpr int lds_floats = 2 * n_age * 64;
  static constexpr int kParA = 5 + n_age;
  static constexpr int kParP64 = 5 + n_age + n_age * n_age;
  static constexpr int kParC = kParP64 + 2;
  static constexpr int kpars = kParC + n_age * n_age;
  template <typename real>
  static constexpr bool coupled = sizeof(real) == 4;

  template <typename real>
  struct shared {
    real beta, p_EI, p_IR, I0;
    double exp_noise;
    const double* kp;  // N, A and C are read from the block as needed
    float a_op[4];     // FP32: this lane's A operand of the 4 k-chunks
  };

  typedef float v4f __attribute__((ext_vector_type(4)));

  template <typename real>
  __device__ __forceinline__ static shared<real> load(const double* kp) {
    shared<real> sh;
    sh.beta = static_cast<real>(kp[0]);
    sh.p_EI = static_cast<real>(sizeof(real) == 8 ? kp[kParP64] : kp[1]);
    sh.p_IR = static_cast<real>(sizeof(real) == 8 ? kp[kParP64 + 1] : kp[2]);
    sh.exp_noise = kp[3];
    sh.I0 = static_cast<real>(kp[4]);
    sh.kp = kp;
    if constexpr (sizeof(real) == 4) {
      const int lane = threadIdx.x % 64;
#pragma unroll
      for (int kk = 0; kk < 4; ++kk) {
        sh.a_op[kk] = static_cast<float>(kp[kParA + (lane % 16) * n_age + 4 * kk + lane / 16]);
      }
    }
    return sh;
  }

  template <typename real>
  __device__ __forceinline__ static int size(const shared<real>&) {
    return n_state;
  }

  template <typename real>
  __device__ __forceinline__ static void initial(const shared<real>& sh, rng256&, bool,
                                                 real* y) {
#pragma unroll
    for (int a = 0; a < n_age; ++a) {
      y[a] = static_cast<real>(sh.kp[5 + a]) - sh.I0;
      y[n_age + a] = 0;
      y[2 * n_age + a] = sh.I0;
      y[3 * n_age + a] = 0;
      y[4 * n_age + a] = 0;
    }
  }

  // FP32: the force of infection of the wavefront's particles into
  // lds[n_age * 64 ..) as [age][lane].  Every lane of the wavefront takes
  // part (lanes without a particle hold zeros).
  __device__ __forceinline__ static void prepare(const float* y, const shared<float>& sh,
                                                 float* lds, int lane) {
    float* xi = lds;
    float* foi = lds + n_age * 64;
#pragma unroll
    for (int a = 0; a < n_age; ++a) xi[a * 64 + lane] = y[2 * n_age + a];
    __syncthreads();
#pragma unroll
    for (int tile = 0; tile < 4; ++tile) {
      v4f d = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int kk = 0; kk < 4; ++kk) {
        const float b = xi[(4 * kk + lane / 16) * 64 + tile * 16 + lane % 16];
        d = __builtin_amdgcn_mfma_f32_16x16x4f32(sh.a_op[kk], b, d, 0, 0, 0);
      }
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        foi[(4 * (lane / 16) + i) * 64 + tile * 16 + lane % 16] = d[i];
      }
    }
    __syncthreads();
  }
  __device__ __forceinline__ static void prepare(const double*, const shared<double>&, float*,
                                                 int) {}

  template <typename real>
  __device__ __forceinline__ static void update(real dt, real* y, const shared<real>& sh,
                                                const float* lds, int lane, rng256& rng,
                                                bool det, uint32_t& err, real& err_a,
                                                real& err_b) {
    real x[n_age];
    if constexpr (sizeof(real) == 8) {
#pragma unroll
      for (int b = 0; b < n_age; ++b) x[b] = y[2 * n_age + b] / sh.kp[5 + b];
    }
#pragma unroll
    for (int a = 0; a < n_age; ++a) {
      real p_SE;
      if constexpr (sizeof(real) == 8) {
        double foi = 0;
        for (int b = 0; b < n_age; ++b) foi += sh.kp[kParC + a * n_age + b] * x[b];
        p_SE = prob_from_rate(-sh.beta * foi * dt);
      } else {
        p_SE = prob_from_rate(-lds[(n_age + a) * 64 + lane] * dt);
      }
      const real S = y[a], E = y[n_age + a], I = y[2 * n_age + a];
      const real n_SE = binomial<real>(rng, S, p_SE, det, err);
      if (err) {
        err_a = S;
        err_b = p_SE;
        return;
      }
      const real n_EI = binomial<real>(rng, E, sh.p_EI, det, err);
      if (err) {
        err_a = E;
        err_b = sh.p_EI;
        return;
      }
      const real n_IR = binomial<real>(rng, I, sh.p_IR, det, err);
      if (err) {
        err_a = I;
        err_b = sh.p_IR;
        return;
      }
      y[a] = S - n_SE;
      y[n_age + a] = E + n_SE - n_EI;
      y[2 * n_age + a] = I + n_EI - n_IR;
      y[3 * n_age + a] += n_IR;
      y[4 * n_age + a] += n_SE;
    }
  }

  template <typename real>
  __device__ __forceinline__ static double compare(const real* y, const double* obs,
                                                   const double* lfact, const shared<real>& sh,
                                                   rng256& rng, bool det) {
    double w = 0;
#pragma unroll
    for (int a = 0; a < n_age; ++a) {
      const double xo = obs[a];
      if (isnan(xo)) continue;
      double noise;
      if (sizeof(real) == 8) {
        noise = exponential_rate<double>(rng, sh.exp_noise, det);
      } else {
        noise = static_cast<double>(
            exponential_rate<float>(rng, static_cast<float>(sh.exp_noise), det));
      }
      w += dpois_log(xo, static_cast<double>(y[4 * n_age + a]) + noise, lfact[a]);
    }
    return w;
  }
};

// Random walk, inst/examples/walk.cpp:1-100: len independent walkers per
// particle, x_i <- Normal(x_i, sd * dt) every step; the initial state is
// Normal(0, sd) per walker when random_initial, else 0.  No data.
// Host parameter row: sd, len, random_initial.  Kernel block: the same.
struct walk_model {
  static constexpr int max_len = 8;
  static constexpr int n_state = max_len;
  static constexpr int n_data = 0;
  static constexpr double zero_period = 1;
  static constexpr int zero_first = 0;
  static constexpr int zero_count = 0;
  static constexpr bool has_beta_schedule = false;
  static constexpr int lds_floats = 0;
  template <typename real>
  static constexpr bool coupled = false;

  template <typename real>
  struct shared {
    real sd;
    int len;
    bool random_initial;
  };

  template <typename real>
  __device__ __forceinline__ static shared<real> load(const double* kp) {
    return shared<real>{static_cast<real>(kp[0]), static_cast<int>(kp[1]), kp[2] != 0};
  }

  template <typename real>
  __device__ __forceinline__ static int size(const shared<real>& sh) {
    return sh.len;
  }

  template <typename real>
  __device__ __forceinline__ static void initial(const shared<real>& sh, rng256& rng, bool det,
                                                 real* y) {
#pragma unroll
    for (int i = 0; i < max_len; ++i) {
      y[i] = 0;
      if (i < sh.len && sh.random_initial) y[i] = normal<real>(rng, 0, sh.sd, det);
    }
  }

  template <typename real>
  __device__ __forceinline__ static void update(real dt, real* y, const shared<real>& sh,
                                                const float*, int, rng256& rng, bool det,
                                                uint32_t&, real&, real&) {
#pragma unroll
    for (int i = 0; i < max_len; ++i) {
      if (i < sh.len) y[i] = normal<real>(rng, y[i], sh.sd * dt, det);
    }
  }

  template <typename real>
  __device__ __forceinline__ static double compare(const real*, const double*, const double*,
                                                   const shared<real>&, rng256&, bool) {
    return 0;
  }
};

// a / b correctly rounded, from inv_b = RN(1 / b): q = RN(a * inv_b) is
// within an ulp of a / b, the residual a - q * b is exact in one fused
// multiply-add, and RN(q + residual * inv_b) is the correctly rounded
// quotient (Markstein's theorem: P. Markstein, "Computation of elementary
// functions on the IBM RISC System/6000 processor", IBM J. Res. Dev. 34
// (1990) 111-119; an inverse within half an ulp of 1 / b suffices).  Three
// dependent operations instead of a full division, for finite a and b in
// the normal range (b = 0 gives NaN where a / b gives an infinity).
template <typename real>
__device__ __forceinline__ real div_rn(real a, real b, real inv_b) {
  const real q = a * inv_b;
  const real r = fma(-q, b, a);
  return fma(r, inv_b, q);
}

// Deterministic SIR ODE, inst/examples/sirode.cpp:9-102.
// Host-side parameter vector per set: N, I0, beta, gamma.
struct sirode_model {
  static constexpr int n_state = 5;
  static constexpr int n_pars = 4;
  static constexpr double zero_period = 1;
  static constexpr int zero_first = 4;
  static constexpr int zero_count = 1;

  template <typename real>
  struct shared {
    real N, I0, beta, gamma;
    real inv_N;  // RN(1 / N): the divisions by N below are div_rn
  };

  template <typename real>
  __device__ __forceinline__ static shared<real> load(const double* pr) {
    shared<real> sh;
    sh.N = static_cast<real>(pr[0]);
    sh.I0 = static_cast<real>(pr[1]);
    sh.beta = static_cast<real>(pr[2]);
    sh.gamma = static_cast<real>(pr[3]);
    sh.inv_N = 1 / sh.N;
    return sh;
  }

  template <typename real>
  __device__ __forceinline__ static void initial(const shared<real>& sh, real* y) {
    y[0] = sh.N - sh.I0;
    y[1] = sh.I0;
    y[2] = 0;
    y[3] = 0;
    y[4] = 0;
  }

  template <typename real>
  __device__ __forceinline__ static void rhs(const real* y, const shared<real>& sh, real* dy) {
    const real S = y[0];
    const real I = y[1];
    const real rate_SI = div_rn(sh.beta * S * I, sh.N, sh.inv_N);
    const real rate_IR = sh.gamma * I;
    dy[0] = -rate_SI;
    dy[1] = rate_SI - rate_IR;
    dy[2] = rate_IR;
    dy[3] = rate_SI;
    dy[4] = rate_SI;
  }

  // Forward sensitivity s = dy/dtheta (theta: 0 beta, 1 gamma, 2 I0):
  // ds/dt = J(y) s + df/dtheta.
  template <typename real>
  __device__ __forceinline__ static void rhs_sens(const real* y, const real* s,
                                                  const shared<real>& sh, int param,
                                                  real* ds) {
    const real S = y[0];
    const real I = y[1];
    real d_SI = div_rn(sh.beta * (s[0] * I + S * s[1]), sh.N, sh.inv_N);
    real d_IR = sh.gamma * s[1];
    if (param == 0) d_SI += div_rn(S * I, sh.N, sh.inv_N);
    if (param == 1) d_IR += I;
    ds[0] = -d_SI;
    ds[1] = d_SI - d_IR;
    ds[2] = d_IR;
    ds[3] = d_SI;
    ds[4] = d_SI;
  }

  // dy0/dtheta of the initial state
  template <typename real>
  __device__ __forceinline__ static void initial_sens(int param, real* s) {
    for (int i = 0; i < 5; ++i) s[i] = 0;
    if (param == 2) {
      s[0] = -1;
      s[1] = 1;
    }
  }
};

}  // namespace dust2_amd
