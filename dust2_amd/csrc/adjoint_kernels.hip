// Deterministic discrete-time SIR unfilter and its reverse-mode (adjoint)
// gradient with respect to beta, gamma and I0.
//
// Replaces unfilter<T>::run / run_adjoint / compute_gradient_
// (dust2/unfilter.hpp:42-114,188-229), the history-saving run and the
// adjoint sweeps of the discrete system (dust2/discrete/system.hpp:94-119,
// 290-379,425-482) and the model's adjoint functions
// (inst/examples/sir.cpp:122-198), for one thread per particle.
//
// Forward: the interval schedule (start time and step count per data time)
// comes from the host; when the gradient is wanted the S and I values that
// enter every step are written to an HBM history [step][2][n_total]
// (coalesced across the particles of a launch) together with the incidence
// at each data time [t][n_total].  Backward: the 8 adjoint values (5 states
// + 3 parameters) stay in registers and walk the history in reverse.
#include <hip/hip_runtime.h>

#include "device_random.hpp"
#include "engine.hpp"
#include "models.hpp"

namespace dust2_amd {

namespace {

// Per particle-thread: k = g * P + p.  Parameter block per group (8
// doubles): N, I0, beta, gamma, exp_noise, p_IR = 1 - exp(-gamma dt),
// e_IR = exp(-gamma dt), exp(-gamma) (all from the host).
__global__ void __launch_bounds__(256) sir_unfilter_kernel(sir_unfilter_launch a) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= a.n_total) return;
  const int g = k / a.n_particles;
  if (a.active != nullptr && !a.active[g]) return;
  const double* pg = a.pars + 8 * g;
  sir_model::shared<double> sh;
  sh.N = pg[0];
  sh.I0 = pg[1];
  sh.beta = pg[2];
  sh.gamma = pg[3];
  sh.exp_noise = pg[4];
  sh.p_IR = pg[5];
  const double e_IR = pg[6];
  const double e_gamma = pg[7];
  const size_t NT = static_cast<size_t>(a.n_total);

  rng256 det_rng{0, 0, 0, 0};
  double y[5];
  if (a.set_initial) {
    sir_model::initial(sh, det_rng, true, y);
  } else {
#pragma unroll
    for (int c = 0; c < 5; ++c) y[c] = a.state[c * NT + k];
  }
  double ll = 0;
  size_t step = 0;
  for (int i = 0; i < a.n_time; ++i) {
    const double t0 = a.sched_time[i];
    const uint32_t ns = a.sched_steps[i];
    for (uint32_t s = 0; s < ns; ++s, ++step) {
      const double time_s = t0 + s * a.dt;
      if (fmod(time_s, sir_model::zero_period) == 0) {
        y[sir_model::zero_first] = 0;
      }
      if (a.hist != nullptr) {
        a.hist[(2 * step) * NT + k] = y[0];
        a.hist[(2 * step + 1) * NT + k] = y[1];
      }
      uint32_t err = 0;
      double ea = 0, eb = 0;
      sir_model::update<double>(a.dt, y, sh, nullptr, 0, det_rng, true, err, ea, eb);
      if (err) {
        a.errors[k] = error_record{err, static_cast<uint32_t>(i), ea, eb};
        atomicAdd(a.error_count, 1u);
        return;
      }
    }
    const size_t o = static_cast<size_t>(i) * a.obs_stride + g * a.obs_group_stride;
    ll += sir_model::compare<double>(y, a.obs + o, a.lfact + o, sh, det_rng, true);
    if (a.inc != nullptr) a.inc[i * NT + k] = y[4];
    if (a.traj != nullptr) {
#pragma unroll
      for (int c = 0; c < 5; ++c) a.traj[(static_cast<size_t>(i) * 5 + c) * NT + k] = y[c];
    }
  }
#pragma unroll
  for (int c = 0; c < 5; ++c) a.state[c * NT + k] = y[c];
  a.ll[k] = ll;
  if (a.grad == nullptr) return;

  // Reverse sweep (dust2/unfilter.hpp:188-229): at each data time, the
  // compare adjoint, then the interval's steps in reverse, re-applying the
  // incidence reset after each reversed step whose start time is on the
  // reset period (dust2/discrete/system.hpp:453-482).
  double adj[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  const double noise = 1 / sh.exp_noise;
  for (int i = a.n_time - 1; i >= 0; --i) {
    const size_t o = static_cast<size_t>(i) * a.obs_stride + g * a.obs_group_stride;
    const double x = a.obs[o];
    // the reference's compare adjoint has no missing-data branch
    // (sir.cpp:179-197), so NaN data give it a NaN gradient
    if (!a.exact || !isnan(x)) {
      adj[4] += x / (a.inc[i * NT + k] + noise) - 1;
    }
    const double t_end = a.sched_time[i] + a.sched_steps[i] * a.dt;
    for (uint32_t s = 1; s <= a.sched_steps[i]; ++s) {
      --step;
      const double S = a.hist[(2 * step) * NT + k];
      const double I = a.hist[(2 * step + 1) * NT + k];
      // sir.cpp:152-176.  Exact: the derivative of p_IR = 1 - exp(-gamma dt)
      // is dt * exp(-gamma dt); the reference writes exp(-gamma) and
      // evaluates p_SI in the forward model's association
      const double e_SI = exp(-sh.beta * I * a.dt / sh.N);
      const double p_SI = a.exact ? 1 - e_SI : 1 - exp(-sh.beta * I / sh.N * a.dt);
      const double adj_n_IR = -adj[1] + adj[2];
      const double adj_n_SI = adj[4] + adj[3] + adj[1] - adj[0];
      const double adj_p_IR = adj_n_IR * I;
      const double adj_p_SI = adj_n_SI * S;
      adj[0] = adj_n_SI * p_SI + adj[0];
      adj[1] = adj_n_IR * sh.p_IR + adj_p_SI * (sh.beta * a.dt) * e_SI / sh.N + adj[1];
      adj[5] = adj[5] + adj_p_SI * I * a.dt * e_SI / sh.N;
      adj[6] = a.exact ? adj[6] + adj_p_IR * a.dt * e_IR : adj[6] + adj_p_IR * e_gamma;
      const double time_s = t_end - s * a.dt;
      if (fmod(time_s, sir_model::zero_period) == 0) {
        adj[sir_model::zero_first] = 0;
      }
    }
  }
  // Initial conditions S0 = N - I0, I0 = I0 (sir.cpp:38-48): dS0/dI0 = -1,
  // dI0/dI0 = 1, and a caller-supplied initial state does not depend on I0.
  // The reference's adjoint_initial adds adj_I only (sir.cpp:126-141), and
  // compute_gradient_ applies it whatever the initial state
  // (dust2/unfilter.hpp:219-222).
  if (a.exact) {
    if (a.set_initial) adj[7] += adj[1] - adj[0];
  } else {
    adj[7] += adj[1];
  }
  a.grad[3 * k + 0] = adj[5];
  a.grad[3 * k + 1] = adj[6];
  a.grad[3 * k + 2] = adj[7];
}

}  // namespace

hipError_t launch_sir_unfilter(const sir_unfilter_launch& args, hipStream_t stream) {
  const int block = 256;
  const int grid = (args.n_total + block - 1) / block;
  hipLaunchKernelGGL(sir_unfilter_kernel, dim3(grid), dim3(block), 0, stream, args);
  return hipGetLastError();
}

}  // namespace dust2_amd
