// General continuous-time systems: Dormand-Prince 5(4) with events found by
// root finding on the dense output, carried ("special") variables that only
// events change, periodic resets, a per-particle history of the dense output
// for delayed values, and critical times; one thread per particle, FP64.
//
// Follows ode::solver (dust2/continuous/solver.hpp:95-262 try_step / step /
// run with apply_events at :359-435, initialise at :264-330) and the
// continuous system's use of it (dust2/continuous/system.hpp:80-106 run_to_time,
// :371-410 rhs_ and output_ with delays, :414-482 initialise_delays_ and
// initialise_solver_), history and delays (dust2/continuous/history.hpp,
// delays.hpp), events (dust2/continuous/events.hpp).  Every floating-point
// operation is in the reference's order, so results agree with the reference
// solver to the last place (save for pow, whose device and host versions can
// differ in the last bit).
//
// Per-particle solver internals live in HBM between calls as a row of FP64
// words (layout in ode_general.hpp); the step-time log, the event log and
// the history ring are separate per-particle arrays.
#include <hip/hip_runtime.h>

#include "block_resample.hpp"
#include "engine.hpp"
#include "ode_general.hpp"
#include "ode_models.hpp"

namespace dust2_amd {

namespace {

constexpr double kEps = 2.220446049250313e-16;  // DBL_EPSILON

namespace tab {
constexpr double C2 = 0.2, C3 = 0.3, C4 = 0.8, C5 = 8.0 / 9.0;
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
}  // namespace tab

template <int N>
struct at_least_one {
  static constexpr int value = N > 0 ? N : 1;
};

// Registers of one particle.
template <class M>
struct gstate {
  static constexpr int NV = M::n_var;
  static constexpr int NY = M::n_var + M::n_special;
  static constexpr int ME = at_least_one<M::max_events>::value;
  double y[NY];
  double dydt[NV];
  double t0, t1, h;  // last accepted step (history_step t0, t1, h)
  double c1[NY], c2[NV], c3[NV], c4[NV], c5[NV];
  double step_size, error;
  double log_fac_old;  // log(max(error, 1e-4)), not stored: rebuilt on load
  double n_steps, n_accepted, n_rejected;
  double last_event[ME];
  double hist_count, hist_first_t0;
  double event_count;
};

// What a particle's rhs needs besides its state: parameters and, for
// delays, the history.
struct gctx {
  const double* p;
  const double* hist;       // this particle's ring [cap][entry words]
  const double* hist_init;  // initial values of the history variables
  uint32_t hist_cap;
  double hist_count;
  double hist_first_t0;
  uint32_t* code;
};

// History slot j holds state variable hist_var<M>(j): the delayed
// variables for a model with delays, every variable otherwise.
template <class M>
__host__ __device__ constexpr int hist_var(int j) {
  if constexpr (M::n_delays > 0) {
    return M::hist_index(j);
  } else {
    return j;
  }
}

template <class M>
struct hist_words {
  static constexpr int NH = M::n_delays > 0 ? at_least_one<M::n_hist>::value
                                            : M::n_var + M::n_special;
  static constexpr int value = 3 + 5 * NH;
};

// history<real_type>::interpolate (dust2/continuous/history.hpp:139-160):
// before the first step (or with no steps) the initial values; otherwise the
// first step whose end is at or after `time`.  The ring keeps the last
// hist_cap steps: a time in a step that has left the ring is an error.
template <class M>
__device__ void history_eval(const gctx& c, double time, int slot, double& out) {
  constexpr int NH = hist_words<M>::NH;
  constexpr int W = hist_words<M>::value;
  const uint64_t count = static_cast<uint64_t>(c.hist_count);
  if (count == 0 || time < c.hist_first_t0) {
    out = c.hist_init[slot];
    return;
  }
  const uint64_t cap = c.hist_cap;
  const uint64_t kept = count < cap ? count : cap;
  const uint64_t first = count - kept;  // step number of the oldest kept
  auto entry = [&](uint64_t j) { return c.hist + ((first + j) % cap) * W; };
  if (first > 0 && time < entry(0)[0]) {
    *c.code = ERR_ODE_HISTORY_SHORT;
    out = 0;
    return;
  }
  // lower_bound on t1
  uint64_t lo = 0, hi = kept;
  while (lo < hi) {
    const uint64_t mid = (lo + hi) / 2;
    if (time > entry(mid)[1]) {
      lo = mid + 1;
    } else {
      hi = mid;
    }
  }
  if (lo == kept) {
    *c.code = ERR_ODE_HISTORY_LOCATE;
    out = 0;
    return;
  }
  const double* e = entry(lo);
  const double u = (time - e[0]) / e[2];
  const double v = 1 - u;
  const double* cf = e + 3;
  const int var = hist_var<M>(slot);
  if (var < M::n_var) {
    out = cf[slot] + u * (cf[NH + slot] + v * (cf[2 * NH + slot] +
                                               u * (cf[3 * NH + slot] + v * cf[4 * NH + slot])));
  } else {
    out = cf[slot];
  }
}

// delays<real_type>::eval (dust2/continuous/delays.hpp:94-101): every
// delay's variables at time - tau, concatenated.
template <class M>
__device__ void delays_eval(const gctx& c, double time, double* d) {
  if constexpr (M::n_delays > 0) {
    int k = 0;
    for (int i = 0; i < M::n_delays; ++i) {
      const double tt = time - M::tau(i, c.p);
      for (int j = 0; j < M::delay_len(i); ++j) history_eval<M>(c, tt, M::delay_slot(i, j), d[k++]);
    }
  }
}

template <class M>
struct delay_buf {
  static constexpr int value = 4;  // every model here delays at most 4 values
};

template <class M>
__device__ __forceinline__ void call_rhs(const gctx& c, double t, const double* y, double* dy) {
  double d[delay_buf<M>::value];
  if constexpr (M::rhs_uses_delays) delays_eval<M>(c, t, d);
  M::rhs(t, y, c.p, d, dy);
}

// solver::try_step (solver.hpp:95-149); y_stage carries the special
// variables of y, as the reference's y_next_ / y_stiff_ do.
template <class M>
__device__ double try_step(const gctx& c, double t, double h, const gstate<M>& s,
                           const ode_control_dev& ctl, double* y_next, double* k2, double* c5) {
  using namespace tab;
  constexpr int NV = gstate<M>::NV;
  constexpr int NY = gstate<M>::NY;
  const double* y = s.y;
  const double* k1 = s.dydt;
  double k3[NV], k4[NV], k5[NV], k6[NV], ys[NY];
  for (int i = NV; i < NY; ++i) {
    y_next[i] = y[i];
    ys[i] = y[i];
  }
  const double t_next = t + h;
  for (int i = 0; i < NV; ++i) y_next[i] = y[i] + h * A21 * k1[i];
  call_rhs<M>(c, t + C2 * h, y_next, k2);
  for (int i = 0; i < NV; ++i) y_next[i] = y[i] + h * (A31 * k1[i] + A32 * k2[i]);
  call_rhs<M>(c, t + C3 * h, y_next, k3);
  for (int i = 0; i < NV; ++i)
    y_next[i] = y[i] + h * (A41 * k1[i] + A42 * k2[i] + A43 * k3[i]);
  call_rhs<M>(c, t + C4 * h, y_next, k4);
  for (int i = 0; i < NV; ++i)
    y_next[i] = y[i] + h * (A51 * k1[i] + A52 * k2[i] + A53 * k3[i] + A54 * k4[i]);
  call_rhs<M>(c, t + C5 * h, y_next, k5);
  for (int i = 0; i < NV; ++i)
    ys[i] = y[i] + h * (A61 * k1[i] + A62 * k2[i] + A63 * k3[i] + A64 * k4[i] + A65 * k5[i]);
  call_rhs<M>(c, t_next, ys, k6);
  for (int i = 0; i < NV; ++i)
    y_next[i] = y[i] + h * (A71 * k1[i] + A73 * k3[i] + A74 * k4[i] + A75 * k5[i] + A76 * k6[i]);
  call_rhs<M>(c, t_next, y_next, k2);
  for (int i = 0; i < NV; ++i)
    c5[i] = h * (D1 * k1[i] + D3 * k3[i] + D4 * k4[i] + D5 * k5[i] + D6 * k6[i] + D7 * k2[i]);
  double err = 0;
  for (int i = 0; i < NV; ++i) {
    const double e = h * (E1 * k1[i] + E3 * k3[i] + E4 * k4[i] + E5 * k5[i] + E6 * k6[i] +
                          E7 * k2[i]);
    const double sk = e / (ctl.atol + ctl.rtol * fmax(fabs(y[i]), fabs(y_next[i])));
    err += sk * sk;
  }
  return sqrt(err / NV);
}

// history_step::interpolate of variable i at t (history.hpp:49-64)
template <class M>
__device__ __forceinline__ double dense(const gstate<M>& s, int i, double t) {
  if (i >= M::n_var) return s.c1[i];
  const double u = (t - s.t0) / s.h;
  const double v = 1 - u;
  return s.c1[i] + u * (s.c2[i] + v * (s.c3[i] + u * (s.c4[i] + v * s.c5[i])));
}

// The model's event function on the dense output of the last step.
template <class M>
__device__ double event_fn(const gstate<M>& s, const double* p, int e, double t) {
  constexpr int NY = gstate<M>::NY;
  double yt[NY];
  for (int i = 0; i < NY; ++i) yt[i] = dense<M>(s, i, t);
  return M::test(e, t, yt, p);
}

// events.hpp:19-44
__device__ __forceinline__ bool is_root(double a, double b, int root) {
  switch (root) {
    case odem::ROOT_INCREASE: return a < 0 && b > 0;
    case odem::ROOT_DECREASE: return a > 0 && b < 0;
    default: return a * b < 0;
  }
}
__device__ __forceinline__ bool is_root_at_b(double a, double b, int root) {
  switch (root) {
    case odem::ROOT_INCREASE: return b == 0 && a < 0;
    case odem::ROOT_DECREASE: return b == 0 && a > 0;
    default: return b == 0 && a != 0;
  }
}

// Brent's root finder on [a, b] where f(a), f(b) have opposite signs
// (R. P. Brent, "Algorithms for Minimization without Derivatives", 1973,
// ch. 4, the zeroin procedure as published on netlib; the reference uses
// the same procedure, inst/include/lostturnip.hpp:29-134): inverse
// quadratic or secant interpolation guarded by bisection, stopping when the
// bracket is within 2 eps |b| + tol / 2; returns the last estimate after
// max_iter evaluations.
template <class F>
__device__ double brent_zero(F f, double a, double b, double tol, int max_iter) {
  double fa = f(a);
  double fb = f(b);
  if (fa == 0) return a;
  if (fb == 0) return b;
  if (fa * fb > 0) return __builtin_nan("");
  double c = a, fc = fa;
  for (int it = 0; it < max_iter; ++it) {
    const double prev_step = b - a;
    if (fabs(fc) < fabs(fb)) {
      a = b;
      b = c;
      c = a;
      fa = fb;
      fb = fc;
      fc = fa;
    }
    const double tol_act = 2 * kEps * fabs(b) + tol / 2;
    double new_step = (c - b) / 2;
    if (fabs(new_step) <= tol_act || fb == 0) return b;
    if (fabs(prev_step) >= tol_act && fabs(fa) > fabs(fb)) {
      double p, q;
      const double cb = c - b;
      if (a == c) {
        const double t1 = fb / fa;
        p = cb * t1;
        q = 1 - t1;
      } else {
        q = fa / fc;
        const double t1 = fb / fc;
        const double t2 = fb / fa;
        p = t2 * (cb * q * (q - t1) - (b - a) * (t1 - 1));
        q = (q - 1) * (t1 - 1) * (t2 - 1);
      }
      if (p > 0) {
        q = -q;
      } else {
        p = -p;
      }
      if (p < (0.75 * cb * q - fabs(tol_act * q) / 2) && p < fabs(prev_step * q / 2)) {
        new_step = p / q;
      }
    }
    if (fabs(new_step) < tol_act) new_step = copysign(tol_act, new_step);
    a = b;
    fa = fb;
    b += new_step;
    fb = f(b);
    if ((fb > 0 && fc > 0) || (fb < 0 && fc < 0)) {
      c = a;
      fc = fa;
    }
  }
  return b;
}

struct glogs {
  double* events;  // [n_total][cap][3]: time, index, sign
  uint32_t event_cap;
  double* hist;    // this particle's ring
  uint32_t hist_cap;
  bool save_history;
  ode_record rec;
  int i;
};

// solver::apply_events (solver.hpp:359-435): looks for a root of each event
// function over the step just taken (skipping an event already found at
// the step start), narrows the step end to the earliest root, then sets
// y_next to the dense output there and applies the actions of every event
// found at that time.  Returns the (possibly earlier) step end.
template <class M>
__device__ bool apply_events(double t0, double h, gstate<M>& s, const double* p, double* y_next,
                             double& t_out, const glogs& lg) {
  constexpr int ME = gstate<M>::ME;
  constexpr int NY = gstate<M>::NY;
  double t1 = t0 + h;
  bool found[ME];
  double sign[ME];
  bool found_any = false;
  const int n_events = M::n_events(p);
  for (int e = 0; e < ME; ++e) found[e] = false;
  for (int e = 0; e < n_events; ++e) {
    if (s.last_event[e] == t0) continue;
    const int root = M::root(e, p);
    const double f_t0 = event_fn<M>(s, p, e, t0);
    const double f_t1 = event_fn<M>(s, p, e, t1);
    if (is_root(f_t0, f_t1, root)) {
      // tolerance 1e-6, at most 100 iterations (solver.hpp:387-389)
      t1 = brent_zero([&](double t) { return event_fn<M>(s, p, e, t); }, t0, t1, 1e-6, 100);
      if (found_any) {
        for (int j = 0; j < ME; ++j) found[j] = false;
      }
    } else if (!is_root_at_b(f_t0, f_t1, root)) {
      continue;
    }
    sign[e] = f_t0 < 0 ? 1 : -1;
    found[e] = true;
    found_any = true;
  }
  if (found_any) {
    for (int i = 0; i < NY; ++i) y_next[i] = dense<M>(s, i, t1);
    for (int e = 0; e < n_events; ++e) {
      if (found[e]) {
        M::action(e, t1, sign[e], y_next, p);
        if (lg.events) {
          const uint64_t n = static_cast<uint64_t>(s.event_count);
          if (n < lg.event_cap) {
            double* r = lg.events + (static_cast<size_t>(lg.i) * lg.event_cap + n) * 3;
            r[0] = t1;
            r[1] = e;
            r[2] = sign[e];
          }
        }
        s.event_count += 1;
        s.last_event[e] = t1;
      }
    }
    s.t1 = t1;
  }
  t_out = t1;
  return found_any;
}

__device__ __forceinline__ void record_time(const ode_record& rec, int i, double t) {
  if (rec.times == nullptr) return;
  const uint32_t n = rec.count[i];
  if (n < rec.cap) rec.times[static_cast<size_t>(i) * rec.cap + n] = t;
  rec.count[i] = n + 1;
}

// internals::save_history (solver.hpp:56-60) into the ring
template <class M>
__device__ void save_history(gstate<M>& s, const glogs& lg) {
  if (!lg.save_history || lg.hist == nullptr) return;
  constexpr int NH = hist_words<M>::NH;
  constexpr int W = hist_words<M>::value;
  const uint64_t n = static_cast<uint64_t>(s.hist_count);
  double* e = lg.hist + (n % lg.hist_cap) * W;
  e[0] = s.t0;
  e[1] = s.t1;
  e[2] = s.h;
  for (int j = 0; j < NH; ++j) {
    const int v = hist_var<M>(j);
    const bool ode = v < M::n_var;
    e[3 + j] = s.c1[v];
    e[3 + NH + j] = ode ? s.c2[v] : 0;
    e[3 + 2 * NH + j] = ode ? s.c3[v] : 0;
    e[3 + 3 * NH + j] = ode ? s.c4[v] : 0;
    e[3 + 4 * NH + j] = ode ? s.c5[v] : 0;
  }
  if (n == 0) s.hist_first_t0 = s.t0;
  s.hist_count += 1;
}

// solver::step (solver.hpp:152-229)
template <class M>
__device__ double step(gctx& c, double t, double t_end, gstate<M>& s, const ode_control_dev& ctl,
                       const glogs& lg) {
  constexpr int NV = gstate<M>::NV;
  constexpr int NY = gstate<M>::NY;
  const double facc1 = 1 / ctl.factor_min;
  const double facc2 = 1 / ctl.factor_max;
  bool reject = false;
  bool truncated = false;
  bool event = false;
  double h = s.step_size;
  double y_next[NY], k2[NV], c5[NV];
  uint32_t& code = *c.code;
  while (true) {
    if (s.n_steps > ctl.max_steps) {
      code = ERR_ODE_MAX_STEPS;
      return t;
    }
    if (h < ctl.step_size_min) {
      code = ERR_ODE_STEP_SMALL;
      return t;
    }
    if (h <= fabs(t) * kEps) {
      code = ERR_ODE_VANISHED;
      return t;
    }
    truncated = t + h > t_end;
    if (truncated) h = t_end - t;
    const double err = try_step<M>(c, t, h, s, ctl, y_next, k2, c5);
    if (code) return t;
    s.n_steps += 1;
    // err^constant, over max(error_old, 1e-4)^beta when accepted
    // (solver.hpp:182, 210-212), as one exp of logs
    const double log_err = log(err);
    const bool accept = err <= 1;
    const double fac11 =
        exp(ctl.constant * log_err - (accept ? ctl.beta * s.log_fac_old : 0.0));
    if (accept) {
      // update_interpolation (solver.hpp:335-352)
      s.t0 = t;
      s.t1 = t + h;
      s.h = h;
      for (int i = 0; i < NV; ++i) {
        const double ydiff = y_next[i] - s.y[i];
        const double bspl = h * s.dydt[i] - ydiff;
        s.c1[i] = s.y[i];
        s.c2[i] = ydiff;
        s.c3[i] = bspl;
        s.c4[i] = -h * k2[i] + ydiff - bspl;
        s.c5[i] = c5[i];
      }
      for (int i = NV; i < NY; ++i) s.c1[i] = s.y[i];
      if constexpr (M::max_events > 0) {
        double t_next;
        if (apply_events<M>(t, h, s, c.p, y_next, t_next, lg)) {
          event = true;
          truncated = false;
          h = t_next - t;
          for (int i = NV; i < NY; ++i) s.y[i] = y_next[i];
          call_rhs<M>(c, t_next, y_next, k2);
          if (code) return t;
        }
      }
      // accept (solver.hpp:354-357)
      for (int i = 0; i < NV; ++i) {
        s.dydt[i] = k2[i];
        s.y[i] = y_next[i];
      }
      s.n_accepted += 1;
      record_time(lg.rec, lg.i, truncated ? t_end : t + h);
      save_history<M>(s, lg);
      if (!truncated && !event) {
        const double fac = fmax(fmin(fac11 / ctl.factor_safe, facc1), facc2);
        const double h_new = h / fac;
        const double h_max = reject ? h : ctl.step_size_max;
        s.step_size = fmin(h_new, h_max);
        s.error = err;
        s.log_fac_old = fmax(log_err, kLogFacMin);
      }
      return truncated ? t_end : t + h;
    }
    reject = true;
    if (s.n_accepted >= 1) s.n_rejected += 1;
    h /= fmin(facc1, fac11 / ctl.factor_safe);
  }
}

// apply_zero_every / apply_zero_every_final (solver.hpp:437-492)
template <class M>
__device__ void zero_every(double t, gstate<M>& s, const double* p) {
  if constexpr (M::n_zero > 0) {
    if (s.h == 0) return;
    for (int z = 0; z < M::n_zero; ++z) {
      const double period = M::zero_period(z, p);
      const int n = static_cast<int>(floor(t / period));
      const int m = static_cast<int>(floor(s.t0 / period));
      if (n > m) {
        const double t_reset = n * period;
        for (int j = 0; j < M::zero_len; ++j) {
          const int v = M::zero_index(z, j);
          if (t_reset == t) {
            s.y[v] = 0;
          } else {
            s.y[v] -= dense<M>(s, v, t_reset);
          }
        }
      }
    }
  }
}

template <class M>
__device__ void zero_every_final(double t, gstate<M>& s, const double* p) {
  if constexpr (M::n_zero > 0) {
    if (s.h == 0) return;
    for (int z = 0; z < M::n_zero; ++z) {
      const double period = M::zero_period(z, p);
      if (s.h > period) {
        const int n = static_cast<int>(ceil(t / period));
        const int m = static_cast<int>(ceil(s.t0 / period));
        if (n > m) {
          const double t_reset = (n - 1) * period;
          for (int j = 0; j < M::zero_len; ++j) {
            const int v = M::zero_index(z, j);
            s.y[v] -= dense<M>(s, v, t_reset);
          }
        }
      }
    }
  }
}

// solver::run (solver.hpp:231-262)
template <class M>
__device__ double run(gctx& c, double t, double t_end, gstate<M>& s, const ode_control_dev& ctl,
                      const glogs& lg) {
  int ic = 0;
  while (ic < ctl.n_critical && ctl.critical_times[ic] <= t) ++ic;
  double t_end_i =
      (ic == ctl.n_critical || ctl.critical_times[ic] >= t_end) ? t_end : ctl.critical_times[ic];
  while (t < t_end && *c.code == 0) {
    zero_every<M>(t, s, c.p);
    t = step<M>(c, t, t_end_i, s, ctl, lg);
    // the history changed: later delayed values see the new step
    c.hist_count = s.hist_count;
    c.hist_first_t0 = s.hist_first_t0;
    if (t >= t_end_i && t < t_end) {
      ++ic;
      t_end_i = (ic == ctl.n_critical || ctl.critical_times[ic] >= t_end)
                    ? t_end
                    : ctl.critical_times[ic];
    }
  }
  if (*c.code == 0) zero_every_final<M>(t, s, c.p);
  return t;
}

// solver::initialise (solver.hpp:264-324) and initialise_events (:326-332);
// the last step's dense output is left as it was, as internals::reset
// leaves it (solver.hpp:62-70).
template <class M>
__device__ void initialise(gctx& c, double t, gstate<M>& s, const ode_control_dev& ctl) {
  constexpr int NV = gstate<M>::NV;
  constexpr int NY = gstate<M>::NY;
  s.step_size = 0;
  s.error = 0;
  s.log_fac_old = kLogFacMin;
  s.n_steps = 0;
  s.n_accepted = 0;
  s.n_rejected = 0;
  double f1[NV], y1[NY];
  for (int i = NV; i < NY; ++i) y1[i] = s.y[i];
  call_rhs<M>(c, t, s.y, s.dydt);
  double norm_f = 0, norm_y = 0;
  for (int i = 0; i < NV; ++i) {
    const double sk = ctl.atol + ctl.rtol * fabs(s.y[i]);
    const double a = s.dydt[i] / sk;
    const double b = s.y[i] / sk;
    norm_f += a * a;
    norm_y += b * b;
  }
  double h = (norm_f <= 1e-10 || norm_y <= 1e-10) ? 1e-6 : sqrt(norm_y / norm_f) * 0.01;
  h = fmin(h, ctl.step_size_max);
  for (int i = 0; i < NV; ++i) y1[i] = s.y[i] + h * s.dydt[i];
  call_rhs<M>(c, t + h, y1, f1);
  double der2 = 0;
  for (int i = 0; i < NV; ++i) {
    const double sk = ctl.atol + ctl.rtol * fabs(s.y[i]);
    const double d = (f1[i] - s.dydt[i]) / sk;
    der2 += d * d;
  }
  der2 = sqrt(der2) / h;
  const double der12 = fmax(fabs(der2), sqrt(norm_f));
  const double h1 = (der12 <= 1e-15) ? fmax(1e-6, fabs(h) * 1e-3) : pow(0.01 / der12, 1.0 / 5);
  h = fmin(100 * fabs(h), h1);
  if (!isfinite(h)) {
    *c.code = ERR_ODE_INIT;
    return;
  }
  h = fmax(h, ctl.step_size_min * 100);
  h = fmin(h, ctl.step_size_max);
  s.step_size = h;
}

// internals::reset then solver::initialise at t (solver.hpp:62-70,
// 264-324): the step-time log restarts at t and the history at the
// current state.
template <class M>
__device__ void reinitialise(gctx& c, double t, gstate<M>& s, const ode_control_dev& ctl,
                             const glogs& lg, double* hist_init) {
  constexpr int NH = hist_words<M>::NH;
  if (lg.rec.times) lg.rec.count[lg.i] = 0;
  record_time(lg.rec, lg.i, t);
  for (int j = 0; j < NH; ++j) hist_init[j] = s.y[hist_var<M>(j)];
  s.hist_count = 0;
  s.hist_first_t0 = 0;
  c.hist_count = 0;
  c.hist_first_t0 = 0;
  initialise<M>(c, t, s, ctl);
}

// Mixed time (dust2/continuous/system.hpp:544-585, run_to_time1 with
// dt > 0), restated step for step: n = (time - time_) / dt steps; step k
// integrates to time_ + k dt (k < n - 1; the last to `time`), then applies
// the model's update with the system's RNG stream as it stands (the
// reference passes the stream by value, so updates do not advance it) and
// re-initialises the solver at time_ if the update changed the state.
template <class M>
__device__ void run_mixed(gctx& c, double time0, double time, double dt, gstate<M>& s,
                          const ode_control_dev& ctl, const glogs& lg, double* hist_init,
                          const rng256& rng, bool det) {
  constexpr int NY = gstate<M>::NY;
  const uint64_t n_steps = static_cast<uint64_t>((time - time0) / dt);
  double t1 = time0;
  for (uint64_t k = 0; k < n_steps && *c.code == 0; ++k) {
    const double t0 = t1;
    t1 = (k == n_steps - 1) ? time : time0 + static_cast<double>(k) * dt;
    run<M>(c, t0, t1, s, ctl, lg);
    if (*c.code) return;
    double y_next[NY];
    for (int i = 0; i < NY; ++i) y_next[i] = s.y[i];
    rng256 r = rng;
    M::update(time, dt, s.y, c.p, r, det, y_next);
    bool same = true;
    for (int i = 0; i < NY; ++i) {
      same = same && (y_next[i] == s.y[i]);
      s.y[i] = y_next[i];
    }
    if (!same) reinitialise<M>(c, time0, s, ctl, lg, hist_init);
  }
}

template <class M>
__device__ void load(gstate<M>& s, const double* w) {
  constexpr gen_layout L = make_gen_layout(M::n_var, M::n_var + M::n_special,
                                           at_least_one<M::max_events>::value,
                                           hist_words<M>::NH);
  for (int i = 0; i < M::n_var; ++i) s.dydt[i] = w[L.dydt + i];
  s.t0 = w[L.t0];
  s.t1 = w[L.t1];
  s.h = w[L.h];
  for (int i = 0; i < gstate<M>::NY; ++i) s.c1[i] = w[L.c1 + i];
  for (int i = 0; i < M::n_var; ++i) {
    s.c2[i] = w[L.c2 + i];
    s.c3[i] = w[L.c3 + i];
    s.c4[i] = w[L.c4 + i];
    s.c5[i] = w[L.c5 + i];
  }
  s.step_size = w[L.step_size];
  s.error = w[L.error];
  s.log_fac_old = fmax(log(s.error), kLogFacMin);
  s.n_steps = w[L.n_steps];
  s.n_accepted = w[L.n_accepted];
  s.n_rejected = w[L.n_rejected];
  for (int e = 0; e < gstate<M>::ME; ++e) s.last_event[e] = w[L.last_event + e];
  s.hist_count = w[L.hist_count];
  s.hist_first_t0 = w[L.hist_first_t0];
  s.event_count = w[L.event_count];
}

template <class M>
__device__ void store(const gstate<M>& s, double* w) {
  constexpr gen_layout L = make_gen_layout(M::n_var, M::n_var + M::n_special,
                                           at_least_one<M::max_events>::value,
                                           hist_words<M>::NH);
  for (int i = 0; i < M::n_var; ++i) w[L.dydt + i] = s.dydt[i];
  w[L.t0] = s.t0;
  w[L.t1] = s.t1;
  w[L.h] = s.h;
  for (int i = 0; i < gstate<M>::NY; ++i) w[L.c1 + i] = s.c1[i];
  for (int i = 0; i < M::n_var; ++i) {
    w[L.c2 + i] = s.c2[i];
    w[L.c3 + i] = s.c3[i];
    w[L.c4 + i] = s.c4[i];
    w[L.c5 + i] = s.c5[i];
  }
  w[L.step_size] = s.step_size;
  w[L.error] = s.error;
  w[L.n_steps] = s.n_steps;
  w[L.n_accepted] = s.n_accepted;
  w[L.n_rejected] = s.n_rejected;
  for (int e = 0; e < gstate<M>::ME; ++e) w[L.last_event + e] = s.last_event[e];
  w[L.hist_count] = s.hist_count;
  w[L.hist_first_t0] = s.hist_first_t0;
  w[L.event_count] = s.event_count;
}

// output_ (dust2/continuous/system.hpp:389-410): output variables from the
// state and the delayed values at the current time
template <class M>
__device__ void compute_output(const gctx& c, double time, double* y_full) {
  if constexpr (M::n_output > 0) {
    double d[delay_buf<M>::value];
    delays_eval<M>(c, time, d);
    M::output(time, y_full, c.p, d);
  }
}

template <class M>
__global__ void __launch_bounds__(256) general_system_kernel(gen_launch a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n_total) return;
  const int g = i / a.n_particles;
  if (a.active && !a.active[g]) return;
  constexpr int NY = gstate<M>::NY;
  constexpr int NS = NY + M::n_output;
  constexpr int W = hist_words<M>::value;
  constexpr int NH = hist_words<M>::NH;
  constexpr gen_layout L = make_gen_layout(M::n_var, NY, at_least_one<M::max_events>::value, NH);
  const size_t NT = static_cast<size_t>(a.n_total);
  uint32_t code = 0;
  gctx c;
  c.p = a.pars + static_cast<size_t>(g) * a.n_pars;
  c.code = &code;
  double* w = a.internals + static_cast<size_t>(i) * a.words;
  c.hist = a.history ? a.history + static_cast<size_t>(i) * a.history_cap * W : nullptr;
  c.hist_init = w + L.hist_init;
  c.hist_cap = a.history_cap;
  glogs lg;
  lg.events = a.events;
  lg.event_cap = a.event_cap;
  lg.hist = a.history ? a.history + static_cast<size_t>(i) * a.history_cap * W : nullptr;
  lg.hist_cap = a.history_cap;
  lg.save_history = a.save_history != 0;
  lg.rec = a.record;
  lg.i = i;

  // delays::step_size_max (delays.hpp:85-92): with delayed values in the
  // rhs no step may be longer than the shortest delay
  ode_control_dev ctl = a.control;
  if constexpr (M::rhs_uses_delays) {
    for (int d = 0; d < M::n_delays; ++d) {
      const double tau = M::tau(d, c.p);
      ctl.step_size_max = ctl.step_size_max < tau ? ctl.step_size_max : tau;
    }
  }
  gstate<M> s;
  double y_full[NS];
  for (int k = 0; k < NS; ++k) y_full[k] = a.state[k * NT + i];
  for (int k = 0; k < NY; ++k) s.y[k] = y_full[k];
  load<M>(s, w);
  if (a.reinit[g]) {
    // initialise_solver_ (system.hpp:444-482): initialise, then
    // initialise_events
    reinitialise<M>(c, a.time, s, ctl, lg, w + L.hist_init);
    for (int e = 0; e < gstate<M>::ME; ++e) s.last_event[e] = -__builtin_inf();
  }
  c.hist_count = s.hist_count;
  c.hist_first_t0 = s.hist_first_t0;
  rng256 rng{};
  if constexpr (M::mixed) {
    if (a.rng) {
      const uint64_t* r = a.rng + 4 * static_cast<size_t>(i);
      rng = rng256{r[0], r[1], r[2], r[3]};
    }
  }
  double time = a.time;
  for (int q = 0; q < a.n_targets && code == 0; ++q) {
    if (M::mixed && a.dt > 0) {
      run_mixed<M>(c, time, a.targets[q], a.dt, s, ctl, lg, w + L.hist_init, rng,
                   a.deterministic != 0);
    } else {
      run<M>(c, time, a.targets[q], s, ctl, lg);
    }
    time = a.targets[q];
    for (int k = 0; k < NY; ++k) y_full[k] = s.y[k];
    compute_output<M>(c, time, y_full);
    if (a.out) {
      double* o = a.out + static_cast<size_t>(q) * NS * NT;
      for (int k = 0; k < NS; ++k) o[k * NT + i] = y_full[k];
    }
  }
  for (int k = 0; k < NY; ++k) y_full[k] = s.y[k];
  for (int k = 0; k < NS; ++k) a.state[k * NT + i] = y_full[k];
  store<M>(s, w);
  if (code) {
    a.errors[i] = error_record{code, 0u, time, 0.0};
    atomicAdd(a.error_count, 1u);
  }
}

// compare_data (dust2/continuous/system.hpp:287-312): the model's log
// density of one data row (per group, or shared) at each particle's state.
template <class M>
__global__ void __launch_bounds__(256) general_compare_kernel(gen_compare_launch a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n_total) return;
  const int g = i / a.n_particles;
  if (a.active && !a.active[g]) return;
  constexpr int NS = M::n_var + M::n_special + M::n_output;
  const size_t NT = static_cast<size_t>(a.n_total);
  double y[NS];
  for (int k = 0; k < NS; ++k) y[k] = a.state[k * NT + i];
  const double* d = a.data + static_cast<size_t>(a.shared_data ? 0 : g) * M::n_data;
  const double v = M::compare(a.time, y, d, a.pars + static_cast<size_t>(g) * a.n_pars);
  a.out[i] = a.accumulate ? a.out[i] + v : v;
}

// One group per workgroup: log-sum-exp of the particles' log-weights added
// to the group's log-likelihood, one uniform from the group's filter stream
// and systematic resampling, indices exact (block_resample.hpp); no
// weights (all exactly 0) leave the identity (dust2/filter.hpp:56-80).
template <int PPT>
__global__ void __launch_bounds__(256) general_resample_kernel(gen_resample_launch a) {
  extern __shared__ double lds[];
  double* scratch = lds;
  double* cum = lds + kScratchDoubles;
  const int g = a.groups ? static_cast<int>(a.groups[blockIdx.x]) : static_cast<int>(blockIdx.x);
  const int P = a.n_particles;
  const int p0 = threadIdx.x * PPT;
  double w[PPT];
  int idx[PPT];
#pragma unroll
  for (int k = 0; k < PPT; ++k) {
    w[k] = p0 + k < P ? a.ll_step[static_cast<size_t>(g) * P + p0 + k] : 0.0;
  }
  uint64_t* fr = a.filter_rng + 8 * static_cast<size_t>(g);  // stream 2g
  auto u_draw = [&]() {
    rng256 r{fr[0], fr[1], fr[2], fr[3]};
    const double u = random_real<double>(r);
    fr[0] = r.s0;
    fr[1] = r.s1;
    fr[2] = r.s2;
    fr[3] = r.s3;
    return u;
  };
  const resample_result res = block_resample<256, PPT>(w, idx, P, cum, scratch, u_draw);
#pragma unroll
  for (int k = 0; k < PPT; ++k) {
    if (p0 + k < P) a.idx[static_cast<size_t>(g) * P + p0 + k] = static_cast<uint32_t>(idx[k]);
  }
  if (threadIdx.x == 0) {
    a.ll[g] += res.ll_increment;
    if (res.n_fallback) atomicAdd(a.fallback_count, 1u);
  }
}

// reorder (dust2/continuous/system.hpp:172-199): state and solver internals
// of slot i come from particle idx[i] of the same group; the particle
// streams stay put.  With traj, the reordered state and the ancestors are
// also saved for time slot t ([G][T][S][P], [G][T][P]).
__global__ void __launch_bounds__(256) general_reorder_kernel(gen_reorder_launch a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n_total) return;
  const int P = a.n_particles;
  const int g = i / P, p = i - g * P;
  const size_t NT = static_cast<size_t>(a.n_total);
  const bool moved = a.active == nullptr || a.active[g];
  const uint32_t from = moved ? static_cast<uint32_t>(g) * P + a.idx[i] : static_cast<uint32_t>(i);
  for (int k = 0; k < a.S; ++k) {
    const double v = a.state_in[k * NT + from];
    a.state_out[k * NT + i] = v;
    if (a.traj) {
      a.traj[((static_cast<size_t>(g) * a.n_time + a.t) * a.S + k) * P + p] = v;
    }
  }
  const double* src = a.words_in + static_cast<size_t>(from) * a.words;
  double* dst = a.words_out + static_cast<size_t>(i) * a.words;
  for (int k = 0; k < a.words; ++k) dst[k] = src[k];
  if (a.ancestry) {
    a.ancestry[(static_cast<size_t>(g) * a.n_time + a.t) * P + p] = moved ? a.idx[i] : p;
  }
}

template <class M>
hipError_t launch_model(const gen_launch& a, hipStream_t stream) {
  const unsigned blocks = static_cast<unsigned>((a.n_total + 255) / 256);
  hipLaunchKernelGGL(general_system_kernel<M>, dim3(blocks), dim3(256), 0, stream, a);
  return hipGetLastError();
}

}  // namespace

int gen_words(int model) {
  const odem::model_info m = odem::info(model);
  const int ny = m.n_var + m.n_special;
  const int nh = m.has_delays ? m.n_hist : ny;
  return make_gen_layout(m.n_var, ny, m.max_events > 0 ? m.max_events : 1, nh).words;
}

int gen_history_words(int model) {
  const odem::model_info m = odem::info(model);
  const int nh = m.has_delays ? m.n_hist : m.n_var + m.n_special;
  return 3 + 5 * nh;
}

hipError_t launch_general_compare(const gen_compare_launch& a, hipStream_t stream) {
  if (a.n_total == 0) return hipSuccess;
  const unsigned blocks = static_cast<unsigned>((a.n_total + 255) / 256);
  switch (a.model) {
    case odem::MODEL_MALARIA:
      hipLaunchKernelGGL(general_compare_kernel<odem::malaria>, dim3(blocks), dim3(256), 0,
                         stream, a);
      return hipGetLastError();
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_general_resample(const gen_resample_launch& a, hipStream_t stream) {
  if (a.n_groups == 0) return hipSuccess;
  const int P = a.n_particles;
  const size_t lds = (kScratchDoubles + static_cast<size_t>(P)) * sizeof(double);
  const dim3 grid(a.n_groups), block(256);
  if (P <= 256) {
    hipLaunchKernelGGL(general_resample_kernel<1>, grid, block, lds, stream, a);
  } else if (P <= 512) {
    hipLaunchKernelGGL(general_resample_kernel<2>, grid, block, lds, stream, a);
  } else if (P <= 1024) {
    hipLaunchKernelGGL(general_resample_kernel<4>, grid, block, lds, stream, a);
  } else if (P <= 2048) {
    hipLaunchKernelGGL(general_resample_kernel<8>, grid, block, lds, stream, a);
  } else if (P <= 4096) {
    hipLaunchKernelGGL(general_resample_kernel<16>, grid, block, lds, stream, a);
  } else if (P <= 8192) {
    hipLaunchKernelGGL(general_resample_kernel<32>, grid, block, lds, stream, a);
  } else {
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_general_reorder(const gen_reorder_launch& a, hipStream_t stream) {
  if (a.n_total == 0) return hipSuccess;
  const unsigned blocks = static_cast<unsigned>((a.n_total + 255) / 256);
  hipLaunchKernelGGL(general_reorder_kernel, dim3(blocks), dim3(256), 0, stream, a);
  return hipGetLastError();
}

hipError_t launch_general_system(const gen_launch& a, hipStream_t stream) {
  if (a.n_total == 0) return hipSuccess;
  if (a.words != gen_words(a.model)) return hipErrorInvalidValue;
  switch (a.model) {
    case odem::MODEL_BOUNCE: return launch_model<odem::bounce>(a, stream);
    case odem::MODEL_CHANGE: return launch_model<odem::change>(a, stream);
    case odem::MODEL_EVENTZERO: return launch_model<odem::eventzero>(a, stream);
    case odem::MODEL_SPECIALEVENT: return launch_model<odem::specialevent>(a, stream);
    case odem::MODEL_SIRDELAY: return launch_model<odem::sirdelay>(a, stream);
    case odem::MODEL_LOGDELAY: return launch_model<odem::logdelay>(a, stream);
    case odem::MODEL_SPECIAL: return launch_model<odem::special>(a, stream);
    case odem::MODEL_MALARIA: return launch_model<odem::malaria>(a, stream);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace dust2_amd
