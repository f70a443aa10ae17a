// Continuous-time models for the general ODE engine (events, carried
// "special" variables, output variables computed from delayed values,
// periodic resets).  Each model is a set of static __host__ __device__
// functions over a flat FP64 parameter row, following the generated-class
// contract of dust2 (initial / rhs / output / events / zero_every / delays;
// dust2/properties.hpp:13-107).  The state of a particle is laid out as the
// reference lays it out (dust2/continuous/system.hpp:44-47): the n_var
// integrated variables, then the n_special carried variables, then the
// n_output variables.
//
// Models (the reference's own test and example systems):
//   bounce        tests/testthat/examples/event.cpp        one root event
//   change        tests/testthat/examples/event-time.cpp   events scheduled in time
//   eventzero     tests/testthat/examples/event-zero.cpp   events without action,
//                                                          a reset of one variable
//   specialevent  tests/testthat/examples/specialevent.cpp an event that switches
//                                                          a carried variable
//   sirdelay      tests/testthat/examples/delay.cpp        incidence from a delayed
//                                                          cumulative count
//   logdelay      a delayed logistic (Hutchinson) equation, whose derivative
//                 reads the delayed state: the rhs side of the delay machinery
//                 (dust2/continuous/system.hpp:371-387, delays.hpp:94-101)
#pragma once

#include <cmath>
#include <cstdint>

namespace dust2_amd {
namespace odem {

// dust2/continuous/events.hpp:12-16
enum root_type : int { ROOT_BOTH = 0, ROOT_INCREASE = 1, ROOT_DECREASE = 2 };

enum model_id : int {
  MODEL_BOUNCE = 0,
  MODEL_CHANGE = 1,
  MODEL_EVENTZERO = 2,
  MODEL_SPECIALEVENT = 3,
  MODEL_SIRDELAY = 4,
  MODEL_LOGDELAY = 5,
  MODEL_COUNT = 6
};

// ---- bounce: a ball under gravity, reflected with damping at height 0 ------
// pars: g, height, velocity, damp
struct bounce {
  static constexpr int id = MODEL_BOUNCE;
  static constexpr int n_var = 2, n_special = 0, n_output = 0;
  static constexpr int n_pars = 4;
  static constexpr int max_events = 1;
  static constexpr int n_zero = 0;
  static constexpr int n_delays = 0, n_hist = 0;
  static constexpr bool rhs_uses_delays = false;
  __host__ __device__ static void initial(double, const double* p, double* y) {
    y[0] = p[1];
    y[1] = p[2];
  }
  __host__ __device__ static void rhs(double, const double* y, const double* p, const double*,
                                      double* dy) {
    dy[0] = y[1];
    dy[1] = -p[0];
  }
  __host__ __device__ static int n_events(const double*) { return 1; }
  __host__ __device__ static int root(int, const double*) { return ROOT_BOTH; }
  __host__ __device__ static double test(int, double, const double* y, const double*) {
    return y[0];
  }
  __host__ __device__ static void action(int, double, double, double* y, const double* p) {
    y[0] = 0;
    y[1] = -p[3] * y[1];
  }
};

// ---- change: constant growth with jumps at scheduled times -----------------
// pars: r, n, t_change[16], delta[16]
struct change {
  static constexpr int id = MODEL_CHANGE;
  static constexpr int max_changes = 16;
  static constexpr int n_var = 1, n_special = 0, n_output = 0;
  static constexpr int n_pars = 2 + 2 * max_changes;
  static constexpr int max_events = max_changes;
  static constexpr int n_zero = 0;
  static constexpr int n_delays = 0, n_hist = 0;
  static constexpr bool rhs_uses_delays = false;
  __host__ __device__ static void initial(double, const double*, double* y) { y[0] = 0; }
  __host__ __device__ static void rhs(double, const double*, const double* p, const double*,
                                      double* dy) {
    dy[0] = p[0];
  }
  __host__ __device__ static int n_events(const double* p) { return static_cast<int>(p[1]); }
  __host__ __device__ static int root(int, const double*) { return ROOT_BOTH; }
  __host__ __device__ static double test(int e, double t, const double*, const double* p) {
    return t - p[2 + e];
  }
  __host__ __device__ static void action(int e, double, double, double* y, const double* p) {
    y[0] += p[2 + max_changes + e];
  }
};

// ---- eventzero: two accumulating variables, the second reset every period,
// with events that only record when each crosses a value ---------------------
// pars: r, value_y, value_z, period_z
struct eventzero {
  static constexpr int id = MODEL_EVENTZERO;
  static constexpr int n_var = 2, n_special = 0, n_output = 0;
  static constexpr int n_pars = 4;
  static constexpr int max_events = 2;
  static constexpr int n_zero = 1;
  static constexpr int n_delays = 0, n_hist = 0;
  static constexpr bool rhs_uses_delays = false;
  __host__ __device__ static void initial(double, const double*, double* y) {
    y[0] = 0;
    y[1] = 1;
  }
  __host__ __device__ static void rhs(double t, const double*, const double* p, const double*,
                                      double* dy) {
    dy[0] = t / p[0];
    dy[1] = t / p[0];
  }
  __host__ __device__ static int n_events(const double*) { return 2; }
  __host__ __device__ static int root(int, const double*) { return ROOT_INCREASE; }
  __host__ __device__ static double test(int e, double, const double* y, const double* p) {
    return e == 0 ? y[0] - p[1] : y[1] - p[2];
  }
  __host__ __device__ static void action(int, double, double, double*, const double*) {}
  // zero_every: {period_z, {1}}
  __host__ __device__ static double zero_period(int, const double* p) { return p[3]; }
  static constexpr int zero_len = 1;
  __host__ __device__ static int zero_index(int, int) { return 1; }
};

// ---- specialevent: growth at rate r0 until an event at time_activate sets
// the carried variable `active`, after which the rate is r1 ------------------
// pars: r0, r1, time_activate (NaN: no event)
struct specialevent {
  static constexpr int id = MODEL_SPECIALEVENT;
  static constexpr int n_var = 1, n_special = 2, n_output = 0;
  static constexpr int n_pars = 3;
  static constexpr int max_events = 1;
  static constexpr int n_zero = 0;
  static constexpr int n_delays = 0, n_hist = 0;
  static constexpr bool rhs_uses_delays = false;
  __host__ __device__ static void initial(double, const double*, double* y) {
    y[0] = 0;
    y[1] = 0;
    y[2] = 0;
  }
  __host__ __device__ static void rhs(double, const double* y, const double* p, const double*,
                                      double* dy) {
    dy[0] = y[1] != 0 ? p[1] : p[0];
  }
  __host__ __device__ static int n_events(const double* p) { return std::isnan(p[2]) ? 0 : 1; }
  __host__ __device__ static int root(int, const double*) { return ROOT_BOTH; }
  __host__ __device__ static double test(int, double t, const double*, const double* p) {
    return t - p[2];
  }
  __host__ __device__ static void action(int, double, double, double* y, const double*) {
    y[1] = 1;
  }
};

// ---- sirdelay: the SIR ODE whose incidence is the cumulative count less its
// value one time unit earlier (an output variable read through a delay) ------
// pars: N, I0, beta, gamma
struct sirdelay {
  static constexpr int id = MODEL_SIRDELAY;
  static constexpr int n_var = 4, n_special = 0, n_output = 1;
  static constexpr int n_pars = 4;
  static constexpr int max_events = 0;
  static constexpr int n_zero = 0;
  // delays<real_type>({{1, {{3, 1}}}}): tau = 1 on variable 3
  static constexpr int n_delays = 1;
  static constexpr int n_hist = 1;  // variables kept in the history
  static constexpr bool rhs_uses_delays = false;
  __host__ __device__ static int hist_index(int) { return 3; }
  __host__ __device__ static double tau(int, const double*) { return 1; }
  // the delayed value of delay d, element j, sits at history slot:
  __host__ __device__ static int delay_slot(int, int) { return 0; }
  __host__ __device__ static int delay_len(int) { return 1; }
  __host__ __device__ static void initial(double, const double* p, double* y) {
    y[0] = p[0] - p[1];
    y[1] = p[1];
    y[2] = 0;
    y[3] = 0;
    y[4] = 0;
  }
  __host__ __device__ static void rhs(double, const double* y, const double* p, const double*,
                                      double* dy) {
    const double rate_SI = p[2] * y[0] * y[1] / p[0];
    const double rate_IR = p[3] * y[1];
    dy[0] = -rate_SI;
    dy[1] = rate_SI - rate_IR;
    dy[2] = rate_IR;
    dy[3] = rate_SI;
  }
  // delayed values of every delay, concatenated: d[0] = cumulative at t - 1
  __host__ __device__ static void output(double, double* y, const double*, const double* d) {
    y[4] = y[3] - d[0];
  }
  __host__ __device__ static int n_events(const double*) { return 0; }
  __host__ __device__ static int root(int, const double*) { return ROOT_BOTH; }
  __host__ __device__ static double test(int, double, const double*, const double*) { return 1; }
  __host__ __device__ static void action(int, double, double, double*, const double*) {}
};

// ---- logdelay: dy/dt = r y(t) (1 - y(t - tau) / K) -------------------------
// pars: r, K, tau, y0
struct logdelay {
  static constexpr int id = MODEL_LOGDELAY;
  static constexpr int n_var = 1, n_special = 0, n_output = 0;
  static constexpr int n_pars = 4;
  static constexpr int max_events = 0;
  static constexpr int n_zero = 0;
  static constexpr int n_delays = 1;
  static constexpr int n_hist = 1;
  static constexpr bool rhs_uses_delays = true;
  __host__ __device__ static int hist_index(int) { return 0; }
  __host__ __device__ static double tau(int, const double* p) { return p[2]; }
  __host__ __device__ static int delay_slot(int, int) { return 0; }
  __host__ __device__ static int delay_len(int) { return 1; }
  __host__ __device__ static void initial(double, const double* p, double* y) { y[0] = p[3]; }
  __host__ __device__ static void rhs(double, const double* y, const double* p, const double* d,
                                      double* dy) {
    dy[0] = p[0] * y[0] * (1 - d[0] / p[1]);
  }
  __host__ __device__ static void output(double, double*, const double*, const double*) {}
  __host__ __device__ static int n_events(const double*) { return 0; }
  __host__ __device__ static int root(int, const double*) { return ROOT_BOTH; }
  __host__ __device__ static double test(int, double, const double*, const double*) { return 1; }
  __host__ __device__ static void action(int, double, double, double*, const double*) {}
};

// Model sizes the host needs without instantiating a kernel.
struct model_info {
  int n_var, n_special, n_output, n_pars, max_events, n_hist;
  bool has_delays;
};

inline model_info info(int id) {
  switch (id) {
    case MODEL_BOUNCE: return {2, 0, 0, bounce::n_pars, 1, 0, false};
    case MODEL_CHANGE: return {1, 0, 0, change::n_pars, change::max_events, 0, false};
    case MODEL_EVENTZERO: return {2, 0, 0, eventzero::n_pars, 2, 0, false};
    case MODEL_SPECIALEVENT: return {1, 2, 0, specialevent::n_pars, 1, 0, false};
    case MODEL_SIRDELAY: return {4, 0, 1, sirdelay::n_pars, 0, 1, true};
    case MODEL_LOGDELAY: return {1, 0, 0, logdelay::n_pars, 0, 1, true};
    default: return {0, 0, 0, 0, 0, 0, false};
  }
}

}  // namespace odem
}  // namespace dust2_amd
