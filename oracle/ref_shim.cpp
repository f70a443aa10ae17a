// ORACLE — TEST INFRASTRUCTURE ONLY.
//
// Thin C wrapper over the parts of the reference engine that compile on
// their own (no monty, no cpp11, no R): dust2/filter_details.hpp,
// dust2/continuous/solver.hpp (+ control/history/events/delays and the
// root finder inst/include/lostturnip.hpp) and
// dust2/interpolate/interpolate.hpp.  Built by
// oracle/Makefile straight from the reference headers into oracle/_ref/;
// used to pin the CPU restatement in oracle.hpp and to write the golden
// fixtures under tests/golden/.  Nothing here is shipped.
#include <dust2/continuous/delays.hpp>
#include <dust2/continuous/solver.hpp>
#include <dust2/filter_details.hpp>
#include <dust2/interpolate/interpolate.hpp>

#include "oracle.hpp"  // xoshiro256+ streams and the normal sampler (monty is absent)

#include <cstdint>
#include <functional>
#include <string>
#include <cstdio>
#include <vector>

extern "C" {

double ref_scale_log_weights(size_t n, double* w) {
  std::vector<double> v(w, w + n);
  const double ret = dust2::details::scale_log_weights<double>(n, v.begin());
  std::copy(v.begin(), v.end(), w);
  return ret;
}

void ref_resample_weight(size_t n, const double* w, double u, uint32_t* idx) {
  std::vector<double> v(w, w + n);
  std::vector<size_t> out(n);
  dust2::details::resample_weight<double>(n, v.cbegin(), u, out.begin());
  for (size_t i = 0; i < n; ++i) idx[i] = static_cast<uint32_t>(out[i]);
}

// The SIRODE model of inst/examples/sirode.cpp integrated by the
// reference solver, driven the way dust_continuous + unfilter drive it
// (dust2/continuous/system.hpp:80-106,444-482,531-542;
// dust2/unfilter.hpp:42-72,162-186): initialise at time_start, then
// solver.run from each data time to the next, zero_every = {{1, {4}}}.
// pars: n_sets x [N, I0, beta, gamma]; data: n_time (shared).
// out_state: n_sets x n_time x 5 (state at each data time);
// out_stats: n_sets x 4 (n_steps, accepted, rejected, step_size).
int ref_sirode_unfilter(const double* pars, size_t n_sets, double time_start,
                        const double* time, size_t n_time, const double* data,
                        const double* ctl_in, double* ll_out,
                        double* out_state, double* out_stats, char* err,
                        size_t errlen) {
  using dust2::ode::internals;
  using dust2::ode::solver;
  const size_t n_var = 5;
  dust2::ode::control<double> ctl(static_cast<size_t>(ctl_in[0]), ctl_in[1], ctl_in[2],
                      std::max(ctl_in[3], std::numeric_limits<double>::epsilon()),
                      ctl_in[4], std::vector<double>(), false, false);
  for (size_t k = 0; k < n_sets; ++k) {
    const double N = pars[4 * k + 0];
    const double I0 = pars[4 * k + 1];
    const double beta = pars[4 * k + 2];
    const double gamma = pars[4 * k + 3];
    auto rhs = [&](double, const double* y, double* dy) {
      const double S = y[0];
      const double I = y[1];
      const double rate_SI = beta * S * I / N;
      const double rate_IR = gamma * I;
      dy[0] = -rate_SI;
      dy[1] = rate_SI - rate_IR;
      dy[2] = rate_IR;
      dy[3] = rate_SI;
      dy[4] = rate_SI;
    };
    solver<double> s(n_var, 0, ctl);
    internals<double> in(n_var, 0, false);
    dust2::zero_every_type<double> zero_every{{1, {4}}};
    dust2::ode::events_type<double> events;
    std::vector<double> y{N - I0, I0, 0, 0, 0};
    double t = time_start;
    double ll = 0;
    try {
      s.initialise(t, y.data(), in, rhs);
      for (size_t i = 0; i < n_time; ++i) {
        s.run(t, time[i], y.data(), zero_every, events, in, rhs);
        t = time[i];
        const double x = data[i];
        if (!std::isnan(x)) {
          const double lambda = y[4];
          ll += (x == 0 && lambda == 0)
                    ? 0
                    : x * std::log(lambda) - lambda - std::lgamma(x + 1);
        }
        if (out_state) {
          std::copy(y.begin(), y.end(), out_state + (k * n_time + i) * n_var);
        }
      }
    } catch (const std::exception& e) {
      if (err && errlen) std::snprintf(err, errlen, "%s", e.what());
      return 1;
    }
    ll_out[k] = ll;
    if (out_stats) {
      out_stats[4 * k + 0] = in.n_steps;
      out_stats[4 * k + 1] = in.n_steps_accepted;
      out_stats[4 * k + 2] = in.n_steps_rejected;
      out_stats[4 * k + 3] = in.step_size;
    }
  }
  return 0;
}

// The SIRODE model of one particle through the reference solver as a
// continuous-time system drives it (dust2/continuous/system.hpp:80-106,
// 444-482, 531-542): initialise at time_start, then solver.run to each
// target in turn, with control.critical_times and
// debug_record_step_times; after target restart_after (if < n_targets) the
// solver is initialised again, as set_state / update_pars make it.  state at each target: out_state[n][5]; stats
// [4]; step times up to cap into step_times, their number in *n_steps.
int ref_sirode_system(const double* pars, double time_start, const double* targets,
                      size_t n_targets, const double* ctl_in, const double* critical,
                      size_t n_critical, size_t restart_after, double* out_state,
                      double* out_stats, double* step_times, size_t cap, size_t* n_steps,
                      char* err, size_t errlen) {
  using dust2::ode::internals;
  using dust2::ode::solver;
  const size_t n_var = 5;
  dust2::ode::control<double> ctl(static_cast<size_t>(ctl_in[0]), ctl_in[1], ctl_in[2],
                      std::max(ctl_in[3], std::numeric_limits<double>::epsilon()),
                      ctl_in[4], std::vector<double>(critical, critical + n_critical), false,
                      true);
  const double N = pars[0], I0 = pars[1], beta = pars[2], gamma = pars[3];
  auto rhs = [&](double, const double* y, double* dy) {
    const double rate_SI = beta * y[0] * y[1] / N;
    const double rate_IR = gamma * y[1];
    dy[0] = -rate_SI;
    dy[1] = rate_SI - rate_IR;
    dy[2] = rate_IR;
    dy[3] = rate_SI;
    dy[4] = rate_SI;
  };
  solver<double> s(n_var, 0, ctl);
  internals<double> in(n_var, 0, false);
  dust2::zero_every_type<double> zero_every{{1, {4}}};
  dust2::ode::events_type<double> events;
  std::vector<double> y{N - I0, I0, 0, 0, 0};
  double t = time_start;
  try {
    s.initialise(t, y.data(), in, rhs);
    for (size_t i = 0; i < n_targets; ++i) {
      s.run(t, targets[i], y.data(), zero_every, events, in, rhs);
      t = targets[i];
      std::copy(y.begin(), y.end(), out_state + i * n_var);
      if (i == restart_after) {
        // set_state / update_pars mark the solver for initialisation
        // (dust2/continuous/system.hpp:444-482): initialise again from here
        s.initialise(t, y.data(), in, rhs);
      }
    }
  } catch (const std::exception& e) {
    if (err && errlen) std::snprintf(err, errlen, "%s", e.what());
    return 1;
  }
  out_stats[0] = in.n_steps;
  out_stats[1] = in.n_steps_accepted;
  out_stats[2] = in.n_steps_rejected;
  out_stats[3] = in.step_size;
  *n_steps = in.step_times.size();
  for (size_t i = 0; i < std::min(cap, in.step_times.size()); ++i) step_times[i] = in.step_times[i];
  return 0;
}

// A generic logistic ODE through the reference solver (closed form
// y = K / (1 + (K/y0 - 1) exp(-r t)), tests/testthat/helper-dust.R:68-73),
// two independent variables with rates r1, r2, no zero_every.
int ref_logistic(double r1, double r2, double K, double y0, const double* time,
                 size_t n_time, const double* ctl_in, double* out) {
  using namespace dust2::ode;
  dust2::ode::control<double> ctl(static_cast<size_t>(ctl_in[0]), ctl_in[1], ctl_in[2],
                      std::max(ctl_in[3], std::numeric_limits<double>::epsilon()),
                      ctl_in[4], std::vector<double>(), false, false);
  auto rhs = [&](double, const double* y, double* dy) {
    dy[0] = r1 * y[0] * (1 - y[0] / K);
    dy[1] = r2 * y[1] * (1 - y[1] / K);
  };
  solver<double> s(2, 0, ctl);
  internals<double> in(2, 0, false);
  dust2::zero_every_type<double> zero_every;
  events_type<double> events;
  std::vector<double> y{y0, y0};
  double t = 0;
  s.initialise(t, y.data(), in, rhs);
  for (size_t i = 0; i < n_time; ++i) {
    s.run(t, time[i], y.data(), zero_every, events, in, rhs);
    t = time[i];
    out[2 * i] = y[0];
    out[2 * i + 1] = y[1];
  }
  return 0;
}

// The reference's interpolators: scalar classes for one series, array
// classes (dimensions n_series) for several; values out[n_z * n_series].
// Returns 0, or 1 with the error text in err.
int ref_interpolate(int kind, const double* t, size_t n_t, const double* y, size_t n_series,
                    const double* z, size_t n_z, double* out, char* err, size_t errlen) {
  try {
    const std::vector<double> tv(t, t + n_t), yv(y, y + n_t * n_series);
    namespace di = dust2::interpolate;
    if (n_series == 1) {
      for (size_t q = 0; q < n_z; ++q) {
        if (kind == 0) out[q] = di::InterpolateConstant<double>(tv, yv, "t", "y").eval(z[q]);
        if (kind == 1) out[q] = di::InterpolateLinear<double>(tv, yv, "t", "y").eval(z[q]);
        if (kind == 2) out[q] = di::InterpolateSpline<double>(tv, yv, "t", "y").eval(z[q]);
      }
    } else {
      dust2::array::dimensions<1> dim{n_series};
      for (size_t q = 0; q < n_z; ++q) {
        double* o = out + q * n_series;
        if (kind == 0) di::InterpolateConstantArray<double, 1>(tv, yv, dim, "t", "y").eval(z[q], o);
        if (kind == 1) di::InterpolateLinearArray<double, 1>(tv, yv, dim, "t", "y").eval(z[q], o);
        if (kind == 2) di::InterpolateSplineArray<double, 1>(tv, yv, dim, "t", "y").eval(z[q], o);
      }
    }
    return 0;
  } catch (const std::exception& e) {
    std::snprintf(err, errlen, "%s", e.what());
    return 1;
  }
}

// One particle of a continuous-time model with events, special variables,
// resets or delays, driven through the reference solver the way
// dust_continuous drives it (dust2/continuous/system.hpp:80-106 run_to_time,
// :371-410 rhs_ / output_, :414-482 initialise_delays_ /
// initialise_solver_): initialise and initialise_events at time_start, then
// solver.run to each target in turn, output variables from the delayed
// values after each.  The models restate the reference's example systems
// (tests/testthat/examples/event.cpp, event-time.cpp, event-zero.cpp,
// specialevent.cpp, delay.cpp) plus a delayed logistic equation; model ids
// and parameter rows as dust2_amd/csrc/ode_models.hpp.
//
// Outputs: out_state[n_targets][S]; stats[4]; step times (cap, *n_steps);
// events[cap_ev][3] = time, index, sign (*n_events); history entries
// [cap_h][3 + 5 H] = t0, t1, h, c1..c5 of the H history variables, oldest
// first (*n_hist).  y0: the initial state, or null for the model's own.
int ref_ode_general(int model, const double* p, const double* y0, double time_start,
                    const double* targets, size_t n_targets, const double* ctl_in,
                    const double* critical, size_t n_critical, int save_history,
                    double* out_state, double* out_stats, double* step_times, size_t cap,
                    size_t* n_steps, double* events, size_t cap_ev, size_t* n_events,
                    double* hist, size_t cap_h, size_t* n_hist, char* err, size_t errlen) {
  using dust2::ode::event;
  using dust2::ode::events_type;
  using dust2::ode::internals;
  using dust2::ode::root_type;
  using dust2::ode::solver;
  size_t n_var = 0, n_special = 0, n_output = 0;
  std::vector<double> y;
  events_type<double> ev;
  dust2::zero_every_type<double> zero_every;
  std::function<void(double, const double*, double*)> rhs_plain;
  bool rhs_delays = false, has_delays = false;
  std::vector<dust2::ode::delays<double>> dl;
  switch (model) {
    case 0: {  // bounce: g, height, velocity, damp
      n_var = 2;
      y = {p[1], p[2]};
      rhs_plain = [p](double, const double* s, double* d) {
        d[0] = s[1];
        d[1] = -p[0];
      };
      ev.push_back(event<double>("bounce", 0, [p](double, double, double* s) {
        s[0] = 0;
        s[1] = -p[3] * s[1];
      }));
      break;
    }
    case 1: {  // change: r, n, t_change[16], delta[16]
      n_var = 1;
      y = {0};
      rhs_plain = [p](double, const double*, double* d) { d[0] = p[0]; };
      for (int i = 0; i < static_cast<int>(p[1]); ++i) {
        ev.push_back(event<double>(
            "event-" + std::to_string(i + 1), std::vector<size_t>{},
            [p, i](double t, const double*) { return t - p[2 + i]; },
            [p, i](double, double, double* s) { s[0] += p[2 + 16 + i]; }));
      }
      break;
    }
    case 2: {  // eventzero: r, value_y, value_z, period_z
      n_var = 2;
      y = {0, 1};
      rhs_plain = [p](double t, const double*, double* d) {
        d[0] = t / p[0];
        d[1] = t / p[0];
      };
      zero_every.push_back({p[3], {1}});
      auto none = [](double, double, double*) {};
      ev.push_back(event<double>("y", std::vector<size_t>{0},
                                 [p](double, const double* s) { return s[0] - p[1]; }, none,
                                 root_type::increase));
      ev.push_back(event<double>("z", std::vector<size_t>{1},
                                 [p](double, const double* s) { return s[0] - p[2]; }, none,
                                 root_type::increase));
      break;
    }
    case 3: {  // specialevent: r0, r1, time_activate
      n_var = 1;
      n_special = 2;
      y = {0, 0, 0};
      rhs_plain = [p](double, const double* s, double* d) { d[0] = s[1] != 0 ? p[1] : p[0]; };
      if (!std::isnan(p[2])) {
        ev.push_back(event<double>(
            "time_activate", std::vector<size_t>{},
            [p](double t, const double*) { return t - p[2]; },
            [](double, double, double* s) { s[1] = 1; }));
      }
      break;
    }
    case 4: {  // sirdelay: N, I0, beta, gamma; output y4 = y3 - y3(t - 1)
      n_var = 4;
      n_output = 1;
      y = {p[0] - p[1], p[1], 0, 0, 0};
      rhs_plain = [p](double, const double* s, double* d) {
        const double rate_SI = p[2] * s[0] * s[1] / p[0];
        const double rate_IR = p[3] * s[1];
        d[0] = -rate_SI;
        d[1] = rate_SI - rate_IR;
        d[2] = rate_IR;
        d[3] = rate_SI;
      };
      dl.push_back(dust2::ode::delays<double>({{1, {{3, 1}}}}));
      has_delays = true;
      break;
    }
    case 5: {  // logdelay: r, K, tau, y0
      n_var = 1;
      y = {p[3]};
      dl.push_back(dust2::ode::delays<double>({{p[2], {{0, 1}}}}));
      has_delays = true;
      rhs_delays = true;
      break;
    }
    default:
      std::snprintf(err, errlen, "unknown model");
      return 1;
  }
  if (y0) std::copy(y0, y0 + y.size(), y.begin());
  const size_t S = y.size();
  dust2::ode::control<double> ctl(static_cast<size_t>(ctl_in[0]), ctl_in[1], ctl_in[2],
                                  std::max(ctl_in[3], std::numeric_limits<double>::epsilon()),
                                  ctl_in[4], std::vector<double>(critical, critical + n_critical),
                                  save_history != 0, true);
  if (has_delays) {
    ctl.step_size_max = dl[0].step_size_max(ctl.step_size_max, rhs_delays);
    ctl.save_history = true;
  }
  solver<double> s(n_var, n_special, ctl);
  internals<double> in(n_var, n_special, ctl.save_history);
  dust2::ode::delay_result_type<double> res;
  if (has_delays) {
    in.history_values.set_index(dl[0].index());
    res = dl[0].result();
  }
  std::function<void(double, const double*, double*)> rhs;
  if (rhs_delays) {
    rhs = [&, p](double t, const double* st, double* d) {
      dl[0].eval(t, in.history_values, res);
      d[0] = p[0] * st[0] * (1 - res[0].data[0] / p[1]);
    };
  } else {
    rhs = rhs_plain;
  }
  double t = time_start;
  try {
    s.initialise(t, y.data(), in, rhs);
    s.initialise_events(ev, in);
    for (size_t i = 0; i < n_targets; ++i) {
      s.run(t, targets[i], y.data(), zero_every, ev, in, rhs);
      t = targets[i];
      if (n_output > 0) {
        dl[0].eval(t, in.history_values, res);
        y[4] = y[3] - res[0].data[0];
      }
      std::copy(y.begin(), y.end(), out_state + i * S);
    }
  } catch (const std::exception& e) {
    if (err && errlen) std::snprintf(err, errlen, "%s", e.what());
    return 1;
  }
  out_stats[0] = in.n_steps;
  out_stats[1] = in.n_steps_accepted;
  out_stats[2] = in.n_steps_rejected;
  out_stats[3] = in.step_size;
  *n_steps = in.step_times.size();
  for (size_t i = 0; i < std::min(cap, in.step_times.size()); ++i) step_times[i] = in.step_times[i];
  *n_events = in.events.size();
  for (size_t i = 0; i < std::min(cap_ev, in.events.size()); ++i) {
    events[3 * i + 0] = in.events[i].time;
    events[3 * i + 1] = static_cast<double>(in.events[i].index);
    events[3 * i + 2] = in.events[i].sign;
  }
  const auto& hd = in.history_values.data();
  *n_hist = hd.size();
  const size_t H = in.history_values.n_state();
  const size_t W = 3 + 5 * H;
  for (size_t i = 0; i < std::min(cap_h, hd.size()); ++i) {
    double* e = hist + i * W;
    e[0] = hd[i].t0;
    e[1] = hd[i].t1;
    e[2] = hd[i].h;
    for (size_t j = 0; j < H; ++j) {
      e[3 + j] = hd[i].c1[j];
      const bool ode = j < hd[i].c2.size();
      e[3 + H + j] = ode ? hd[i].c2[j] : 0;
      e[3 + 2 * H + j] = ode ? hd[i].c3[j] : 0;
      e[3 + 3 * H + j] = ode ? hd[i].c4[j] : 0;
      e[3 + 4 * H + j] = ode ? hd[i].c5[j] : 0;
    }
  }
  return 0;
}

// A mixed-time model (the reference's special and malaria examples,
// tests/testthat/examples/special.cpp, inst/examples/malaria.cpp) through the
// reference solver, with the system's mixed-time loop restated from
// dust2/continuous/system.hpp:544-585 (run_to_time1 with dt > 0) and the
// model's update drawing from a copy of the particle's stream (the
// reference passes the stream by value).  Model ids and parameter rows as
// dust2_amd/csrc/ode_models.hpp; updates use the oracle's normal sampler
// (monty is not in the reference tree).  rng: the particle's stream, 4
// words.  dt = 0: continuous time only.
int ref_ode_mixed(int model, const double* p, const double* y0, double time_start, double dt,
                  const double* targets, size_t n_targets, const uint64_t* rng,
                  int deterministic, double* out_state, double* out_stats, double* step_times,
                  size_t cap, size_t* n_steps, char* err, size_t errlen) {
  using dust2::ode::internals;
  using dust2::ode::solver;
  size_t n_var = 0;
  std::vector<double> y;
  std::function<void(double, const double*, double*)> rhs;
  std::function<void(const double*, oracle::rng_state&, double*)> update;
  if (model == 6) {  // special: sd
    n_var = 1;
    y = {0, 0};
    rhs = [](double, const double* s, double* d) { d[0] = s[1]; };
    update = [p](const double* s, oracle::rng_state& r, double* next) {
      next[1] = oracle::normal(r, s[1], p[0]);
    };
  } else if (model == 7) {  // malaria
    const int n_rates = static_cast<int>(p[3]);
    n_var = 4 + n_rates;
    y.assign(n_var + 1, 0.0);
    y[0] = 1 - p[7];
    y[1] = p[7];
    y[2] = p[9];
    y[3] = p[8];
    y[4 + n_rates] = p[4];
    rhs = [p, n_rates](double, const double* s, double* d) {
      const double a = p[0], bh = p[1], bv = p[2], n = p[3], mu = p[4], r = p[5], tau = p[6];
      const double Sh = s[0], Ih = s[1], Sv = s[2], Iv = s[3];
      const double* Ev = s + 4;
      const double beta = s[4 + n_rates];
      const double foi_h = a * bh * Iv;
      const double foi_v = a * bv * Ih;
      d[4] = foi_v * Sv - (n / tau) * Ev[0] - mu * Ev[0];
      for (int i = 1; i < n_rates; ++i) {
        d[4 + i] = (n / tau) * Ev[i - 1] - (n / tau) * Ev[i] - mu * Ev[i];
      }
      d[0] = -foi_h * Sh + r * Ih;
      d[1] = foi_h * Sh - r * Ih;
      d[2] = beta * p[9] - foi_v * Sv - mu * Sv;
      d[3] = (n / tau) * Ev[n_rates - 1] - mu * Iv;
    };
    update = [p, n_rates](const double* s, oracle::rng_state& r, double* next) {
      const int i = 4 + n_rates;
      next[i] = s[i] * std::exp(oracle::normal(r, 0, p[10]));
    };
  } else {
    std::snprintf(err, errlen, "unknown model");
    return 1;
  }
  if (y0) std::copy(y0, y0 + y.size(), y.begin());
  const size_t n_special = 1, NY = y.size();
  dust2::ode::control<double> ctl;
  ctl.debug_record_step_times = true;
  solver<double> s(n_var, n_special, ctl);
  internals<double> in(n_var, n_special, false);
  dust2::zero_every_type<double> zero_every;
  dust2::ode::events_type<double> events;
  oracle::rng_state stream{{rng[0], rng[1], rng[2], rng[3]}, deterministic != 0};
  double time_ = time_start;
  try {
    s.initialise(time_, y.data(), in, rhs);
    s.initialise_events(events, in);
    for (size_t q = 0; q < n_targets; ++q) {
      const double time = targets[q];
      if (dt == 0) {
        s.run(time_, time, y.data(), zero_every, events, in, rhs);
      } else {
        const size_t steps = (time - time_) / dt;
        double t1 = time_;
        for (size_t k = 0; k < steps; ++k) {
          const double t0 = t1;
          t1 = (k == steps - 1) ? time : time_ + k * dt;
          s.run(t0, t1, y.data(), zero_every, events, in, rhs);
          std::vector<double> y_other(y.begin(), y.begin() + NY);
          oracle::rng_state r = stream;
          update(y.data(), r, y_other.data());
          const bool same = std::equal(y.begin(), y.begin() + NY, y_other.begin());
          std::copy(y_other.begin(), y_other.end(), y.begin());
          if (!same) s.initialise(time_, y.data(), in, rhs);
        }
      }
      time_ = time;
      std::copy(y.begin(), y.end(), out_state + q * NY);
    }
  } catch (const std::exception& e) {
    if (err && errlen) std::snprintf(err, errlen, "%s", e.what());
    return 1;
  }
  out_stats[0] = in.n_steps;
  out_stats[1] = in.n_steps_accepted;
  out_stats[2] = in.n_steps_rejected;
  out_stats[3] = in.step_size;
  *n_steps = in.step_times.size();
  for (size_t i = 0; i < std::min(cap, in.step_times.size()); ++i) step_times[i] = in.step_times[i];
  return 0;
}

// One group of the malaria particle filter (tests/testthat/test-filter.R:
// 705-722) end to end: P particles through the reference solver with the
// mixed-time loop restated as ref_ode_mixed does, compare_data as
// inst/examples/malaria.cpp:176-188 (binomial log density), then the
// reference's own scale_log_weights and resample_weight
// (dust2/filter_details.hpp) with one uniform from the group's filter
// stream, and the reorder of state and solver internals
// (dust2/filter.hpp:42-104, dust2/continuous/system.hpp:172-199).
// part_rng[P][4]; filt_rng[4] (stream 2g, advanced in place);
// data[n_time][2] = tested, positive; out: *ll, final_state[P][S].
int ref_malaria_filter(const double* p, size_t P, double time_start, const double* times,
                       size_t n_time, const double* data, double dt, const uint64_t* part_rng,
                       uint64_t* filt_rng, double* ll_out, double* final_state, char* err,
                       size_t errlen) {
  using dust2::ode::internals;
  using dust2::ode::solver;
  const int n_rates = static_cast<int>(p[3]);
  const size_t n_var = 4 + n_rates, NY = n_var + 1;
  auto rhs = [p, n_rates](double, const double* s, double* d) {
    const double a = p[0], bh = p[1], bv = p[2], n = p[3], mu = p[4], r = p[5], tau = p[6];
    const double Sh = s[0], Ih = s[1], Sv = s[2], Iv = s[3];
    const double* Ev = s + 4;
    const double beta = s[4 + n_rates];
    const double foi_h = a * bh * Iv;
    const double foi_v = a * bv * Ih;
    d[4] = foi_v * Sv - (n / tau) * Ev[0] - mu * Ev[0];
    for (int i = 1; i < n_rates; ++i) {
      d[4 + i] = (n / tau) * Ev[i - 1] - (n / tau) * Ev[i] - mu * Ev[i];
    }
    d[0] = -foi_h * Sh + r * Ih;
    d[1] = foi_h * Sh - r * Ih;
    d[2] = beta * p[9] - foi_v * Sv - mu * Sv;
    d[3] = (n / tau) * Ev[n_rates - 1] - mu * Iv;
  };
  std::vector<double> y0(NY, 0.0);
  y0[0] = 1 - p[7];
  y0[1] = p[7];
  y0[2] = p[9];
  y0[3] = p[8];
  y0[4 + n_rates] = p[4];
  dust2::ode::control<double> ctl;
  solver<double> s(n_var, 1, ctl);
  std::vector<internals<double>> in(P, internals<double>(n_var, 1, false));
  std::vector<std::vector<double>> y(P, y0);
  dust2::zero_every_type<double> zero_every;
  dust2::ode::events_type<double> events;
  oracle::rng_state fr{{filt_rng[0], filt_rng[1], filt_rng[2], filt_rng[3]}, false};
  double ll = 0;
  double time_ = time_start;
  try {
    for (size_t k = 0; k < P; ++k) {
      s.initialise(time_, y[k].data(), in[k], rhs);
      s.initialise_events(events, in[k]);
    }
    std::vector<double> w(P);
    std::vector<size_t> idx(P);
    for (size_t q = 0; q < n_time; ++q) {
      const double time = times[q];
      for (size_t k = 0; k < P; ++k) {
        if (dt == 0) {
          s.run(time_, time, y[k].data(), zero_every, events, in[k], rhs);
          continue;
        }
        const size_t steps = (time - time_) / dt;
        double t1 = time_;
        for (size_t j = 0; j < steps; ++j) {
          const double t0 = t1;
          t1 = (j == steps - 1) ? time : time_ + j * dt;
          s.run(t0, t1, y[k].data(), zero_every, events, in[k], rhs);
          std::vector<double> y_other(y[k]);
          oracle::rng_state r{{part_rng[4 * k], part_rng[4 * k + 1], part_rng[4 * k + 2],
                               part_rng[4 * k + 3]},
                              false};
          const int i = 4 + n_rates;
          y_other[i] = y[k][i] * std::exp(oracle::normal(r, 0, p[10]));
          const bool same = std::equal(y[k].begin(), y[k].end(), y_other.begin());
          y[k] = y_other;
          if (!same) s.initialise(time_, y[k].data(), in[k], rhs);
        }
      }
      time_ = time;
      const double tested = data[2 * q], positive = data[2 * q + 1];
      bool any = false;
      for (size_t k = 0; k < P; ++k) {
        double v = 0;
        if (!std::isnan(tested) && !std::isnan(positive)) {
          const double prob = y[k][1];
          if (prob == 0) {
            v = positive == 0 ? 0 : -std::numeric_limits<double>::infinity();
          } else if (prob == 1) {
            v = positive == tested ? 0 : -std::numeric_limits<double>::infinity();
          } else {
            v = std::lgamma(tested + 1) - std::lgamma(positive + 1) -
                std::lgamma(tested - positive + 1) + positive * std::log(prob) +
                (tested - positive) * std::log1p(-prob);
          }
        }
        w[k] = v;
        any = any || v != 0;
      }
      if (any) {
        ll += dust2::details::scale_log_weights<double>(P, w.begin());
        const double u = oracle::random_real(fr);
        dust2::details::resample_weight(P, w.begin(), u, idx.begin());
      } else {
        for (size_t k = 0; k < P; ++k) idx[k] = k;
      }
      std::vector<std::vector<double>> y2(P);
      std::vector<internals<double>> in2(in);
      for (size_t k = 0; k < P; ++k) {
        y2[k] = y[idx[k]];
        in2[k] = in[idx[k]];
      }
      y.swap(y2);
      in.swap(in2);
    }
  } catch (const std::exception& e) {
    if (err && errlen) std::snprintf(err, errlen, "%s", e.what());
    return 1;
  }
  *ll_out = ll;
  for (size_t k = 0; k < P; ++k) std::copy(y[k].begin(), y[k].end(), final_state + k * NY);
  std::copy(fr.s, fr.s + 4, filt_rng);
  return 0;
}

}  // extern "C"
