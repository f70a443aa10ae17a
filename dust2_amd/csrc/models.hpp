// Models compiled as device functions, written to dust2's generated-class
// contract: initial / update / compare_data / rhs with shared (per group,
// read-only) parameters and a zero_every reset rule.  An odin2-generated
// class maps onto this shape with its build_shared/build_data kept on the
// host (dust2/properties.hpp:13-107 detects the same members).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "device_random.hpp"

namespace dust2_amd {

// Stochastic SIR, inst/examples/sir.cpp:12-199.
// Host-side parameter vector per group: N, I0, beta, gamma, exp_noise.
struct sir_model {
  static constexpr int n_state = 5;
  static constexpr int n_pars = 5;
  static constexpr int n_data = 1;
  // zero_every = {{1, {4}}}: daily incidence reset (sir.cpp:95-97)
  static constexpr double zero_period = 1;
  static constexpr int zero_first = 4;
  static constexpr int zero_count = 1;

  template <typename real>
  struct shared {
    real N, I0, beta, gamma;
    real p_IR;         // 1 - exp(-gamma*dt), computed on the host
    double exp_noise;  // rate of the compare noise
  };

  template <typename real>
  __device__ __forceinline__ static void initial(const shared<real>& sh, real* y) {
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
                                                rng256& rng, bool det, uint32_t& err,
                                                real& err_a, real& err_b) {
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
  };

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
    const real rate_SI = sh.beta * S * I / sh.N;
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
    real d_SI = sh.beta * (s[0] * I + S * s[1]) / sh.N;
    real d_IR = sh.gamma * s[1];
    if (param == 0) d_SI += S * I / sh.N;
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
