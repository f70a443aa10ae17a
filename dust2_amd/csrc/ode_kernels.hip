// Deterministic likelihood ("unfilter") of an ODE model integrated with
// Dormand-Prince 5(4), Shampine dense output and a PI step-size
// controller, one thread per particle, every data interval in one launch.
//
// Replaces unfilter<dust_continuous<T>>::run (dust2/unfilter.hpp:42-72) with
// dust_continuous::run_to_time / initialise_solver_ / compare_data
// (dust2/continuous/system.hpp:80-106, 287-312, 444-482, 531-542) and
// ode::solver::{initialise, run, step, try_step, apply_zero_every[_final]}
// (dust2/continuous/solver.hpp:95-324, 437-492); tableau and controller
// constants as dust2/continuous/solver.hpp:507-544 and control.hpp:11-24.
//
// `real` is the type of the state and stages.  The error norm, the step
// size, time and the controller are FP64 in both instantiations; with
// real = double every operation follows the CPU order exactly.
#include <hip/hip_runtime.h>

#include "engine.hpp"
#include "models.hpp"

namespace dust2_amd {

namespace {

constexpr int NV = sirode_model::n_state;

// Butcher tableau (Dormand & Prince 1980) and Shampine (1986) dense output.
// The stage times c_i are not needed: the model's rhs is autonomous.
namespace dp {
constexpr double A21 = 0.2;
constexpr double A31 = 3.0 / 40.0, A32 = 9.0 / 40.0;
constexpr double A41 = 44.0 / 45.0, A42 = -56.0 / 15.0, A43 = 32.0 / 9.0;
constexpr double A51 = 19372.0 / 6561.0, A52 = -25360.0 / 2187.0, A53 = 64448.0 / 6561.0,
                 A54 = -212.0 / 729.0;
constexpr double A61 = 9017.0 / 3168.0, A62 = -355.0 / 33.0, A63 = 46732.0 / 5247.0,
                 A64 = 49.0 / 176.0, A65 = -5103.0 / 18656.0;
constexpr double A71 = 35.0 / 384.0, A73 = 500.0 / 1113.0, A74 = 125.0 / 192.0,
                 A75 = -2187.0 / 6784.0, A76 = 11.0 / 84.0;
constexpr double E1 = 71.0 / 57600.0, E3 = -71.0 / 16695.0, E4 = 71.0 / 1920.0,
                 E5 = -17253.0 / 339200.0, E6 = 22.0 / 525.0, E7 = -1.0 / 40.0;
constexpr double D1 = -12715105075.0 / 11282082432.0, D3 = 87487479700.0 / 32700410799.0,
                 D4 = -10690763975.0 / 1880347072.0, D5 = 701980252875.0 / 199316789632.0,
                 D6 = -1453857185.0 / 822651844.0, D7 = 69997945.0 / 29380423.0;
}  // namespace dp

// Per-particle integrator state.  Only the reset variable's dense-output
// coefficients are kept: they are the only ones apply_zero_every reads.
template <typename real>
struct ode_state {
  real y[NV];
  real dydt[NV];
  double step_size;
  double error;
  double log_fac_old;  // log(max(error, 1e-4)), the controller's memory
  double t0, h;  // last accepted step
  real c1, c2, c3, c4, c5;  // dense output of variable zero_first
  uint32_t n_steps, n_accepted, n_rejected;
  // forward sensitivity run: s = dy/dtheta, its derivative and the dense
  // output of its reset component
  real sy[NV];
  real dsydt[NV];
  real sc1, sc2, sc3, sc4, sc5;
  int param;
};

template <typename real>
__device__ __forceinline__ real interp5(double t0, double h, real c1, real c2, real c3, real c4,
                                        real c5, double t) {
  const double u = (t - t0) / h;
  const double v = 1 - u;
  return static_cast<real>(static_cast<double>(c1) +
                           u * (static_cast<double>(c2) +
                                v * (static_cast<double>(c3) +
                                     u * (static_cast<double>(c4) +
                                          v * static_cast<double>(c5)))));
}

template <typename real>
__device__ __forceinline__ real interp(const ode_state<real>& s, double t) {
  return interp5(s.t0, s.h, s.c1, s.c2, s.c3, s.c4, s.c5, t);
}

template <typename real>
__device__ __forceinline__ real interp_sens(const ode_state<real>& s, double t) {
  return interp5(s.t0, s.h, s.sc1, s.sc2, s.sc3, s.sc4, s.sc5, t);
}

template <typename real>
__device__ __forceinline__ void rhs(const real* y, const sirode_model::shared<real>& sh,
                                    real* dy) {
  sirode_model::rhs<real>(y, sh, dy);
}

// One attempted step of size h from (t, y) with k1 = dydt; returns the
// FP64 error norm and leaves the 5th-order solution in y_next, the FSAL
// derivative in k2 and the dense-output coefficient in c5.  With SENS the
// sensitivity s = dy/dtheta takes the same stages (its stage derivatives
// evaluated at the state's stage values) into sy_next / ks2 / sc5; it does
// not enter the error norm, so the step sequence is the plain run's.
template <typename real, bool SENS>
__device__ __forceinline__ double try_step(double t, double h, const ode_state<real>& s,
                                           const sirode_model::shared<real>& sh,
                                           const ode_control_dev& ctl, real* y_next, real* k2,
                                           real& c5_out, real* sy_next, real* ks2,
                                           real& sc5_out) {
  using namespace dp;
  const real* y = s.y;
  const real* k1 = s.dydt;
  const real* sy = s.sy;
  const real* q1 = s.dsydt;
  const real hr = static_cast<real>(h);
  real k3[NV], k4[NV], k5[NV], k6[NV], ys[NV];
  real q3[NV], q4[NV], q5[NV], q6[NV], qs[NV];
  (void)t;
  for (int i = 0; i < NV; ++i) y_next[i] = y[i] + hr * static_cast<real>(A21) * k1[i];
  rhs<real>(y_next, sh, k2);
  if constexpr (SENS) {
    for (int i = 0; i < NV; ++i) sy_next[i] = sy[i] + hr * static_cast<real>(A21) * q1[i];
    sirode_model::rhs_sens<real>(y_next, sy_next, sh, s.param, ks2);
  }
  for (int i = 0; i < NV; ++i)
    y_next[i] = y[i] + hr * (static_cast<real>(A31) * k1[i] + static_cast<real>(A32) * k2[i]);
  rhs<real>(y_next, sh, k3);
  if constexpr (SENS) {
    for (int i = 0; i < NV; ++i)
      sy_next[i] = sy[i] + hr * (static_cast<real>(A31) * q1[i] + static_cast<real>(A32) * ks2[i]);
    sirode_model::rhs_sens<real>(y_next, sy_next, sh, s.param, q3);
  }
  for (int i = 0; i < NV; ++i)
    y_next[i] = y[i] + hr * (static_cast<real>(A41) * k1[i] + static_cast<real>(A42) * k2[i] +
                             static_cast<real>(A43) * k3[i]);
  rhs<real>(y_next, sh, k4);
  if constexpr (SENS) {
    for (int i = 0; i < NV; ++i)
      sy_next[i] = sy[i] + hr * (static_cast<real>(A41) * q1[i] + static_cast<real>(A42) * ks2[i] +
                                 static_cast<real>(A43) * q3[i]);
    sirode_model::rhs_sens<real>(y_next, sy_next, sh, s.param, q4);
  }
  for (int i = 0; i < NV; ++i)
    y_next[i] = y[i] + hr * (static_cast<real>(A51) * k1[i] + static_cast<real>(A52) * k2[i] +
                             static_cast<real>(A53) * k3[i] + static_cast<real>(A54) * k4[i]);
  rhs<real>(y_next, sh, k5);
  if constexpr (SENS) {
    for (int i = 0; i < NV; ++i)
      sy_next[i] = sy[i] + hr * (static_cast<real>(A51) * q1[i] + static_cast<real>(A52) * ks2[i] +
                                 static_cast<real>(A53) * q3[i] + static_cast<real>(A54) * q4[i]);
    sirode_model::rhs_sens<real>(y_next, sy_next, sh, s.param, q5);
  }
  for (int i = 0; i < NV; ++i)
    ys[i] = y[i] + hr * (static_cast<real>(A61) * k1[i] + static_cast<real>(A62) * k2[i] +
                         static_cast<real>(A63) * k3[i] + static_cast<real>(A64) * k4[i] +
                         static_cast<real>(A65) * k5[i]);
  rhs<real>(ys, sh, k6);
  if constexpr (SENS) {
    for (int i = 0; i < NV; ++i)
      qs[i] = sy[i] + hr * (static_cast<real>(A61) * q1[i] + static_cast<real>(A62) * ks2[i] +
                            static_cast<real>(A63) * q3[i] + static_cast<real>(A64) * q4[i] +
                            static_cast<real>(A65) * q5[i]);
    sirode_model::rhs_sens<real>(ys, qs, sh, s.param, q6);
  }
  for (int i = 0; i < NV; ++i)
    y_next[i] = y[i] + hr * (static_cast<real>(A71) * k1[i] + static_cast<real>(A73) * k3[i] +
                             static_cast<real>(A74) * k4[i] + static_cast<real>(A75) * k5[i] +
                             static_cast<real>(A76) * k6[i]);
  rhs<real>(y_next, sh, k2);
  if constexpr (SENS) {
    for (int i = 0; i < NV; ++i)
      sy_next[i] = sy[i] + hr * (static_cast<real>(A71) * q1[i] + static_cast<real>(A73) * q3[i] +
                                 static_cast<real>(A74) * q4[i] + static_cast<real>(A75) * q5[i] +
                                 static_cast<real>(A76) * q6[i]);
    sirode_model::rhs_sens<real>(y_next, sy_next, sh, s.param, ks2);
  }
  {
    const int i = sirode_model::zero_first;
    c5_out = hr * (static_cast<real>(D1) * k1[i] + static_cast<real>(D3) * k3[i] +
                   static_cast<real>(D4) * k4[i] + static_cast<real>(D5) * k5[i] +
                   static_cast<real>(D6) * k6[i] + static_cast<real>(D7) * k2[i]);
    if constexpr (SENS) {
      sc5_out = hr * (static_cast<real>(D1) * q1[i] + static_cast<real>(D3) * q3[i] +
                      static_cast<real>(D4) * q4[i] + static_cast<real>(D5) * q5[i] +
                      static_cast<real>(D6) * q6[i] + static_cast<real>(D7) * ks2[i]);
    }
  }
  double err = 0;
  for (int i = 0; i < NV; ++i) {
    const real e = hr * (static_cast<real>(E1) * k1[i] + static_cast<real>(E3) * k3[i] +
                         static_cast<real>(E4) * k4[i] + static_cast<real>(E5) * k5[i] +
                         static_cast<real>(E6) * k6[i] + static_cast<real>(E7) * k2[i]);
    const double sk =
        static_cast<double>(e) /
        (ctl.atol + ctl.rtol * fmax(fabs(static_cast<double>(y[i])),
                                    fabs(static_cast<double>(y_next[i]))));
    err += sk * sk;
  }
  static_assert(NV == 5, "the mean below divides by 5");
  return sqrt(div_rn(err, 5.0, 0.2));
}

// solver::step: returns the new time, or NaN with *code set on failure.
template <typename real, bool SENS>
__device__ double ode_step(double t, double t_end, ode_state<real>& s,
                           const sirode_model::shared<real>& sh, const ode_control_dev& ctl,
                           uint32_t& code) {
  const double facc1 = 1 / ctl.factor_min;
  const double facc2 = 1 / ctl.factor_max;
  const double inv_safe = 1 / ctl.factor_safe;
  bool reject = false;
  bool truncated = false;
  double h = s.step_size;
  real y_next[NV], k2[NV];
  real sy_next[NV], ks2[NV];
  while (true) {
    if (s.n_steps > ctl.max_steps) {
      code = ERR_ODE_MAX_STEPS;
      return t;
    }
    if (h < ctl.step_size_min) {
      code = ERR_ODE_STEP_SMALL;
      return t;
    }
    if (h <= fabs(t) * 2.220446049250313e-16) {
      code = ERR_ODE_VANISHED;
      return t;
    }
    truncated = t + h > t_end;
    if (truncated) {
      h = t_end - t;
    }
    real c5, sc5 = 0;
    const double err =
        try_step<real, SENS>(t, h, s, sh, ctl, y_next, k2, c5, sy_next, ks2, sc5);
    s.n_steps++;
    // The controller's err^constant, over max(error_old, 1e-4)^beta when
    // the step is accepted (solver.hpp:182, 210-212), as one exp of logs: the
    // log of the old error is kept from its own step.  log(0) = -inf and
    // NaN give the same factors as pow.
    const double log_err = log(err);
    const bool accept = err <= 1;
    const double fac11 =
        exp(ctl.constant * log_err - (accept ? ctl.beta * s.log_fac_old : 0.0));
    if (accept) {
      // dense output + accept (solver.hpp:335-357)
      const int z = sirode_model::zero_first;
      const real ydiff = y_next[z] - s.y[z];
      const real bspl = static_cast<real>(h) * s.dydt[z] - ydiff;
      s.c1 = s.y[z];
      s.c2 = ydiff;
      s.c3 = bspl;
      s.c4 = -static_cast<real>(h) * k2[z] + ydiff - bspl;
      s.c5 = c5;
      if constexpr (SENS) {
        const real sdiff = sy_next[z] - s.sy[z];
        const real sbspl = static_cast<real>(h) * s.dsydt[z] - sdiff;
        s.sc1 = s.sy[z];
        s.sc2 = sdiff;
        s.sc3 = sbspl;
        s.sc4 = -static_cast<real>(h) * ks2[z] + sdiff - sbspl;
        s.sc5 = sc5;
        for (int i = 0; i < NV; ++i) {
          s.dsydt[i] = ks2[i];
          s.sy[i] = sy_next[i];
        }
      }
      s.t0 = t;
      s.h = h;
      for (int i = 0; i < NV; ++i) {
        s.dydt[i] = k2[i];
        s.y[i] = y_next[i];
      }
      s.n_accepted++;
      if (!truncated) {
        const double fac = fmax(fmin(div_rn(fac11, ctl.factor_safe, inv_safe), facc1), facc2);
        const double h_new = h / fac;
        const double h_max = reject ? h : ctl.step_size_max;
        s.step_size = fmin(h_new, h_max);
        s.error = err;
        s.log_fac_old = fmax(log_err, kLogFacMin);
      }
      return truncated ? t_end : t + h;
    }
    reject = true;
    if (s.n_accepted >= 1) {
      s.n_rejected++;
    }
    h /= fmin(facc1, div_rn(fac11, ctl.factor_safe, inv_safe));
  }
}

__device__ __forceinline__ void record_step(const ode_record& rec, int i, double t) {
  if (rec.times == nullptr) return;
  const uint32_t n = rec.count[i];
  if (n < rec.cap) rec.times[static_cast<size_t>(i) * rec.cap + n] = t;
  rec.count[i] = n + 1;
}

// solver::run (solver.hpp:231-262): integrate from t to t_end, stopping at
// every critical time in between, applying the reset of the incidence
// variable at each step start that crossed a reset boundary
// (apply_zero_every, solver.hpp:437-465) and after the last step
// (apply_zero_every_final, solver.hpp:468-492).  Returns the time reached.
template <typename real, bool SENS>
__device__ double ode_run(double t, double t_end, ode_state<real>& s,
                          const sirode_model::shared<real>& sh, const ode_control_dev& ctl,
                          uint32_t& code, const ode_record& rec, int i) {
  // first critical time after t (std::upper_bound)
  int ic = 0;
  while (ic < ctl.n_critical && ctl.critical_times[ic] <= t) ++ic;
  double t_end_i =
      (ic == ctl.n_critical || ctl.critical_times[ic] >= t_end) ? t_end : ctl.critical_times[ic];
  while (t < t_end && code == 0) {
    if (s.h != 0) {
      const double period = sirode_model::zero_period;
      const int n = static_cast<int>(floor(t / period));
      const int m = static_cast<int>(floor(s.t0 / period));
      if (n > m) {
        const double t_reset = n * period;
        if (t_reset == t) {
          s.y[sirode_model::zero_first] = 0;
          if constexpr (SENS) s.sy[sirode_model::zero_first] = 0;
        } else {
          s.y[sirode_model::zero_first] -= interp(s, t_reset);
          if constexpr (SENS) s.sy[sirode_model::zero_first] -= interp_sens(s, t_reset);
        }
      }
    }
    t = ode_step<real, SENS>(t, t_end_i, s, sh, ctl, code);
    if (code == 0) record_step(rec, i, t);
    if (t >= t_end_i && t < t_end) {
      ++ic;
      t_end_i = (ic == ctl.n_critical || ctl.critical_times[ic] >= t_end)
                    ? t_end
                    : ctl.critical_times[ic];
    }
  }
  if (code == 0 && s.h != 0 && s.h > sirode_model::zero_period) {
    const double period = sirode_model::zero_period;
    const int n = static_cast<int>(ceil(t / period));
    const int m = static_cast<int>(ceil(s.t0 / period));
    if (n > m) {
      s.y[sirode_model::zero_first] -= interp(s, (n - 1) * period);
      if constexpr (SENS) s.sy[sirode_model::zero_first] -= interp_sens(s, (n - 1) * period);
    }
  }
  return t;
}

// solver::initialise (solver.hpp:264-324), Hairer's starting step.
template <typename real>
__device__ void ode_initialise(ode_state<real>& s, const sirode_model::shared<real>& sh,
                               const ode_control_dev& ctl, uint32_t& code) {
  s.step_size = 0;
  s.error = 0;
  s.log_fac_old = kLogFacMin;
  s.n_steps = 0;
  s.n_accepted = 0;
  s.n_rejected = 0;
  real f1[NV], y1[NV];
  rhs<real>(s.y, sh, s.dydt);
  double norm_f = 0, norm_y = 0;
  for (int i = 0; i < NV; ++i) {
    const double sk = ctl.atol + ctl.rtol * fabs(static_cast<double>(s.y[i]));
    const double a = static_cast<double>(s.dydt[i]) / sk;
    const double b = static_cast<double>(s.y[i]) / sk;
    norm_f += a * a;
    norm_y += b * b;
  }
  double h = (norm_f <= 1e-10 || norm_y <= 1e-10) ? 1e-6 : sqrt(norm_y / norm_f) * 0.01;
  h = fmin(h, ctl.step_size_max);
  for (int i = 0; i < NV; ++i) y1[i] = s.y[i] + static_cast<real>(h) * s.dydt[i];
  rhs<real>(y1, sh, f1);
  double der2 = 0;
  for (int i = 0; i < NV; ++i) {
    const double sk = ctl.atol + ctl.rtol * fabs(static_cast<double>(s.y[i]));
    const double d = (static_cast<double>(f1[i]) - static_cast<double>(s.dydt[i])) / sk;
    der2 += d * d;
  }
  der2 = sqrt(der2) / h;
  const double der12 = fmax(fabs(der2), sqrt(norm_f));
  const double h1 = (der12 <= 1e-15) ? fmax(1e-6, fabs(h) * 1e-3) : pow(0.01 / der12, 1.0 / 5);
  h = fmin(100 * fabs(h), h1);
  if (!isfinite(h)) {
    code = ERR_ODE_INIT;
    return;
  }
  h = fmax(h, ctl.step_size_min * 100);
  h = fmin(h, ctl.step_size_max);
  s.step_size = h;
}

template <typename real, bool SENS>
__global__ void __launch_bounds__(256) dp5_unfilter_kernel(ode_launch args) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= args.n_total) return;
  const int g = i / args.n_particles;
  if (args.active && !args.active[g]) return;
  const double* pr = args.pars + 4 * g;
  const sirode_model::shared<real> sh = sirode_model::load<real>(pr);
  const ode_control_dev ctl = args.control;
  real* state = static_cast<real*>(args.state);

  ode_state<real> s;
  sirode_model::initial<real>(sh, s.y);
  for (int k = 0; k < NV; ++k) s.dydt[k] = 0;
  s.t0 = 0;
  s.h = 0;
  s.c1 = s.c2 = s.c3 = s.c4 = s.c5 = 0;
  uint32_t code = 0;
  ode_initialise<real>(s, sh, ctl, code);
  double dll = 0;
  if constexpr (SENS) {
    s.param = args.sens_param;
    sirode_model::initial_sens<real>(s.param, s.sy);
    sirode_model::rhs_sens<real>(s.y, s.sy, sh, s.param, s.dsydt);
    s.sc1 = s.sc2 = s.sc3 = s.sc4 = s.sc5 = 0;
  }
  double time = args.time_start;
  double ll = 0;
  size_t dstride = args.obs_stride;
  size_t doff = args.obs_group_stride * static_cast<size_t>(g);
  const double* times = args.times;
  int n_time = args.n_time;
  if (args.series_offset) {
    const int64_t off = args.series_offset[g];
    times += off;
    doff = static_cast<size_t>(off);
    dstride = 1;
    n_time = args.series_len[g];
    if (args.series_time_start) time = args.series_time_start[g];
  }
  const ode_record no_record{nullptr, nullptr, 0};
  for (int t = 0; t < n_time && code == 0; ++t) {
    const double t_end = times[t];
    ode_run<real, SENS>(time, t_end, s, sh, ctl, code, no_record, i);
    time = t_end;
    if (args.traj != nullptr) {
      real* tr = static_cast<real*>(args.traj) + static_cast<size_t>(t) * NV * args.n_total;
      for (int k = 0; k < NV; ++k) tr[static_cast<size_t>(k) * args.n_total + i] = s.y[k];
    }
    // compare: log Poisson(obs; cases_inc) (sirode.cpp:89-101)
    const double x = args.obs[t * dstride + doff];
    if (!isnan(x)) {
      const double lambda = static_cast<double>(s.y[4]);
      double v;
      if (x == 0 && lambda == 0) {
        v = 0;
      } else if (lambda == 0) {
        v = -__builtin_inf();
      } else {
        v = x * log(lambda) - lambda - args.lfact[t * dstride + doff];
      }
      ll += v;
      if constexpr (SENS) {
        // d/dlambda log Poisson(x; lambda) = x / lambda - 1
        dll += (x == 0 ? -1.0 : x / lambda - 1) * static_cast<double>(s.sy[4]);
      }
    }
  }
  args.ll[i] = ll;
  for (int k = 0; k < NV; ++k) state[static_cast<size_t>(k) * args.n_total + i] = s.y[k];
  if (args.stats) {
    args.stats[4 * static_cast<size_t>(i) + 0] = s.n_steps;
    args.stats[4 * static_cast<size_t>(i) + 1] = s.n_accepted;
    args.stats[4 * static_cast<size_t>(i) + 2] = s.n_rejected;
    args.stats[4 * static_cast<size_t>(i) + 3] = s.step_size;
  }
  if constexpr (SENS) {
    args.grad[3 * static_cast<size_t>(i) + args.sens_param] = dll;
  }
  if (code) {
    args.errors[i] = error_record{code, 0u, 0.0, 0.0};
    *args.error_flag = 1u;
  }
}

// The continuous-time system (dust2/continuous/system.hpp:80-106,
// 444-482, 531-542): the solver internals of each particle persist in HBM
// between calls; a group whose state, time or parameters changed
// initialises its solver first (initialise_solver_), recording the
// starting time among the step times as solver::initialise does.
__global__ void __launch_bounds__(256) dp5_system_kernel(ode_sys_launch a) {
  static_assert(NV == 5, "ode_internals holds five variables");
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n_total) return;
  const int g = i / a.n_particles;
  const double* pr = a.pars + 4 * g;
  const sirode_model::shared<double> sh = sirode_model::load<double>(pr);
  const size_t NT = static_cast<size_t>(a.n_total);
  ode_state<double> s;
  for (int k = 0; k < NV; ++k) s.y[k] = a.state[k * NT + i];
  uint32_t code = 0;
  ode_internals& in = a.internals[i];
  // the last step's dense output survives initialisation (internals::reset,
  // solver.hpp:62-70, leaves it), so a reset boundary crossed before a
  // set_state still applies at the next run's start
  s.t0 = in.t0;
  s.h = in.h;
  s.c1 = in.c1;
  s.c2 = in.c2;
  s.c3 = in.c3;
  s.c4 = in.c4;
  s.c5 = in.c5;
  if (a.reinit[g]) {
    if (a.record.times) a.record.count[i] = 0;
    record_step(a.record, i, a.time);
    ode_initialise<double>(s, sh, a.control, code);
  } else {
    for (int k = 0; k < NV; ++k) s.dydt[k] = in.dydt[k];
    s.step_size = in.step_size;
    s.error = in.error;
    s.log_fac_old = fmax(log(in.error), kLogFacMin);
    s.n_steps = in.n_steps;
    s.n_accepted = in.n_accepted;
    s.n_rejected = in.n_rejected;
  }
  double time = a.time;
  for (int q = 0; q < a.n_targets && code == 0; ++q) {
    ode_run<double, false>(time, a.targets[q], s, sh, a.control, code, a.record, i);
    time = a.targets[q];
    if (a.out) {
      double* o = a.out + static_cast<size_t>(q) * NV * NT;
      for (int k = 0; k < NV; ++k) o[k * NT + i] = s.y[k];
    }
  }
  for (int k = 0; k < NV; ++k) a.state[k * NT + i] = s.y[k];
  for (int k = 0; k < NV; ++k) in.dydt[k] = s.dydt[k];
  in.step_size = s.step_size;
  in.error = s.error;
  in.t0 = s.t0;
  in.h = s.h;
  in.c1 = s.c1;
  in.c2 = s.c2;
  in.c3 = s.c3;
  in.c4 = s.c4;
  in.c5 = s.c5;
  in.n_steps = s.n_steps;
  in.n_accepted = s.n_accepted;
  in.n_rejected = s.n_rejected;
  if (code) {
    a.errors[i] = error_record{code, 0u, 0.0, 0.0};
    atomicAdd(a.error_count, 1u);
  }
}

}  // namespace

hipError_t launch_dp5_system(const ode_sys_launch& args, hipStream_t stream) {
  if (args.n_total == 0) return hipSuccess;
  const unsigned blocks = static_cast<unsigned>((args.n_total + 255) / 256);
  hipLaunchKernelGGL(dp5_system_kernel, dim3(blocks), dim3(256), 0, stream, args);
  return hipGetLastError();
}

hipError_t launch_dp5_unfilter(const ode_launch& args, bool fp64, hipStream_t stream) {
  if (args.n_total == 0) return hipSuccess;
  const unsigned blocks = static_cast<unsigned>((args.n_total + 255) / 256);
  if (args.sens_param >= 0) {
    if (!fp64 || args.grad == nullptr || args.sens_param > 2) return hipErrorInvalidValue;
    hipLaunchKernelGGL((dp5_unfilter_kernel<double, true>), dim3(blocks), dim3(256), 0, stream,
                       args);
  } else if (fp64) {
    hipLaunchKernelGGL((dp5_unfilter_kernel<double, false>), dim3(blocks), dim3(256), 0,
                       stream, args);
  } else {
    hipLaunchKernelGGL((dp5_unfilter_kernel<float, false>), dim3(blocks), dim3(256), 0,
                       stream, args);
  }
  return hipGetLastError();
}

}  // namespace dust2_amd
