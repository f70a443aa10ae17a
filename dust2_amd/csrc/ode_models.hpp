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
//   special       tests/testthat/examples/special.cpp      mixed time: a carried
//                                                          rate updated by a normal
//                                                          random walk every dt
//   malaria       inst/examples/malaria.cpp                mixed time: host-vector
//                                                          malaria with a random-walk
//                                                          biting rate; compare to
//                                                          binomial prevalence data
//
// Models with an `update` run in mixed time when the system has dt > 0
// (dust2/properties.hpp:99, dust2/continuous/system.hpp:544-585).
#pragma once

#include <cmath>
#include <cstdint>

#include "device_random.hpp"

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
  MODEL_SPECIAL = 6,
  MODEL_MALARIA = 7,
  MODEL_COUNT = 8
};

// Models without a discrete-time update or a data comparison.
struct no_update {
  static constexpr bool mixed = false;
  static constexpr bool has_compare = false;
  static constexpr int n_data = 0;
  __host__ __device__ static void update(double, double, const double*, const double*, rng256&,
                                         bool, double*) {}
  __host__ __device__ static double compare(double, const double*, const double*,
                                            const double*) {
    return 0;
  }
};

// log binomial probability of x successes in `size` trials
// (monty::density::binomial as malaria.cpp:186 calls it; R's dbinom(log = TRUE))
__host__ __device__ inline double binomial_log_density(double x, double size, double prob) {
  if (prob == 0) return x == 0 ? 0 : -INFINITY;
  if (prob == 1) return x == size ? 0 : -INFINITY;
  if (x == 0 && size == 0) return 0;
  const double lchoose = lgamma(size + 1) - lgamma(x + 1) - lgamma(size - x + 1);
  return lchoose + x * log(prob) + (size - x) * log1p(-prob);
}

// ---- bounce: a ball under gravity, reflected with damping at height 0 ------
// pars: g, height, velocity, damp
struct bounce : no_update {
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
struct change : no_update {
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
struct eventzero : no_update {
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
struct specialevent : no_update {
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
struct sirdelay : no_update {
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
struct logdelay : no_update {
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

// ---- special: a = integral of a carried rate b, b a normal random walk ------
// pars: sd
struct special {
  static constexpr int id = MODEL_SPECIAL;
  static constexpr int n_var = 1, n_special = 1, n_output = 0;
  static constexpr int n_pars = 1;
  static constexpr int max_events = 0;
  static constexpr int n_zero = 0;
  static constexpr int n_delays = 0, n_hist = 0;
  static constexpr bool rhs_uses_delays = false;
  static constexpr bool mixed = true;
  static constexpr bool has_compare = false;
  static constexpr int n_data = 0;
  __host__ __device__ static void initial(double, const double*, double* y) {
    y[0] = 0;
    y[1] = 0;
  }
  __host__ __device__ static void rhs(double, const double* y, const double*, const double*,
                                      double* dy) {
    dy[0] = y[1];
  }
  __device__ static void update(double, double, const double* y, const double* p, rng256& rng,
                                bool det, double* y_next) {
    y_next[1] = normal<double>(rng, y[1], p[0], det);
  }
  __host__ __device__ static double compare(double, const double*, const double*,
                                            const double*) {
    return 0;
  }
  __host__ __device__ static int n_events(const double*) { return 0; }
  __host__ __device__ static int root(int, const double*) { return ROOT_BOTH; }
  __host__ __device__ static double test(int, double, const double*, const double*) { return 1; }
  __host__ __device__ static void action(int, double, double, double*, const double*) {}
};

// ---- malaria: humans (Sh, Ih), vectors (Sv, Iv) and n_rates exposed-vector
// stages, with the vector birth rate `beta` a carried variable following a
// log-normal random walk every dt ------------------------------------------
// pars: a, bh, bv, n_rates, mu, r, tau, initial_Ih, initial_Iv, initial_Sv,
//       beta_volatility (n_rates fixed at 15 in this build)
struct malaria {
  static constexpr int id = MODEL_MALARIA;
  static constexpr int n_rates = 15;
  static constexpr int n_var = 4 + n_rates, n_special = 1, n_output = 0;
  static constexpr int n_pars = 11;
  static constexpr int max_events = 0;
  static constexpr int n_zero = 0;
  static constexpr int n_delays = 0, n_hist = 0;
  static constexpr bool rhs_uses_delays = false;
  static constexpr bool mixed = true;
  static constexpr bool has_compare = true;
  static constexpr int n_data = 2;  // tested, positive
  __host__ __device__ static void initial(double, const double* p, double* y) {
    y[0] = 1 - p[7];
    y[1] = p[7];
    y[2] = p[9];
    y[3] = p[8];
    for (int i = 4; i < 4 + n_rates; ++i) y[i] = 0;
    y[4 + n_rates] = p[4];
  }
  __host__ __device__ static void rhs(double, const double* y, const double* p, const double*,
                                      double* dy) {
    const double a = p[0], bh = p[1], bv = p[2], n = p[3], mu = p[4], r = p[5], tau = p[6];
    const double Sh = y[0], Ih = y[1], Sv = y[2], Iv = y[3];
    const double* Ev = y + 4;
    const double beta = y[4 + n_rates];
    const double foi_h = a * bh * Iv;
    const double foi_v = a * bv * Ih;
    dy[4] = foi_v * Sv - (n / tau) * Ev[0] - mu * Ev[0];
    for (int i = 1; i < n_rates; ++i) {
      dy[4 + i] = (n / tau) * Ev[i - 1] - (n / tau) * Ev[i] - mu * Ev[i];
    }
    dy[0] = -foi_h * Sh + r * Ih;
    dy[1] = foi_h * Sh - r * Ih;
    dy[2] = beta * p[9] - foi_v * Sv - mu * Sv;
    dy[3] = (n / tau) * Ev[n_rates - 1] - mu * Iv;
  }
  __device__ static void update(double, double, const double* y, const double* p, rng256& rng,
                                bool det, double* y_next) {
    const int i = 4 + n_rates;
    y_next[i] = y[i] * exp(normal<double>(rng, 0, p[10], det));
  }
  // data: tested, positive (malaria.cpp:176-188)
  __host__ __device__ static double compare(double, const double* y, const double* d,
                                            const double*) {
    if (std::isnan(d[1]) || std::isnan(d[0])) return 0;
    return binomial_log_density(d[1], d[0], y[1]);
  }
  __host__ __device__ static int n_events(const double*) { return 0; }
  __host__ __device__ static int root(int, const double*) { return ROOT_BOTH; }
  __host__ __device__ static double test(int, double, const double*, const double*) { return 1; }
  __host__ __device__ static void action(int, double, double, double*, const double*) {}
};

// Model sizes the host needs without instantiating a kernel.
struct model_info {
  int n_var, n_special, n_output, n_pars, max_events, n_hist;
  bool has_delays;
  bool mixed;
  int n_data;
};

inline model_info info(int id) {
  switch (id) {
    case MODEL_BOUNCE: return {2, 0, 0, bounce::n_pars, 1, 0, false, false, 0};
    case MODEL_CHANGE: return {1, 0, 0, change::n_pars, change::max_events, 0, false, false, 0};
    case MODEL_EVENTZERO: return {2, 0, 0, eventzero::n_pars, 2, 0, false, false, 0};
    case MODEL_SPECIALEVENT: return {1, 2, 0, specialevent::n_pars, 1, 0, false, false, 0};
    case MODEL_SIRDELAY: return {4, 0, 1, sirdelay::n_pars, 0, 1, true, false, 0};
    case MODEL_LOGDELAY: return {1, 0, 0, logdelay::n_pars, 0, 1, true, false, 0};
    case MODEL_SPECIAL: return {1, 1, 0, special::n_pars, 0, 0, false, true, 0};
    case MODEL_MALARIA:
      return {malaria::n_var, 1, 0, malaria::n_pars, 0, 0, false, true, malaria::n_data};
    default: return {0, 0, 0, 0, 0, 0, false, false, 0};
  }
}

}  // namespace odem
}  // namespace dust2_amd
