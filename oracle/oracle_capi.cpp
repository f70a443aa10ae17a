// ORACLE — TEST INFRASTRUCTURE ONLY (see oracle.hpp).  C entry points over
// the CPU restatement, loaded by tests/ (ctypes), the smoke check and the
// CPU-baseline leg of bench.py.  Never linked into the product.
#include "oracle.hpp"

#include <chrono>
#include <cstring>

using namespace oracle;

namespace {

int fail(const std::exception& e, char* err, size_t errlen) {
  if (err && errlen) {
    std::snprintf(err, errlen, "%s", e.what());
  }
  return 1;
}

std::vector<size_t> make_index(const size_t* index_group, size_t n_index,
                               size_t n_groups) {
  std::vector<size_t> ret;
  if (index_group == nullptr) {
    ret.resize(n_groups);
    std::iota(ret.begin(), ret.end(), 0);
  } else {
    ret.assign(index_group, index_group + n_index);
  }
  return ret;
}

template <typename T>
std::vector<typename T::shared_state> read_shared(const double* pars,
                                                  size_t n_groups) {
  std::vector<typename T::shared_state> ret;
  for (size_t g = 0; g < n_groups; ++g) {
    ret.push_back(T::build_shared(pars + g * T::n_pars));
  }
  return ret;
}

template <typename T>
std::vector<typename T::data_type> read_data(const double* data, size_t n_rows) {
  std::vector<typename T::data_type> ret;
  for (size_t r = 0; r < n_rows; ++r) {
    ret.push_back(T::build_data(data + r * T::n_data));
  }
  return ret;
}

// Expand the G*(P+2) filter streams from an integer seed or take them
// verbatim, and split them [2 filter | P system] per group
// (dust2/r/discrete/filter.hpp:63-74).
void split_filter_streams(const std::vector<rng_state>& all, size_t G, size_t P,
                          std::vector<uint64_t>& seed_filter,
                          std::vector<uint64_t>& seed_system) {
  seed_filter.clear();
  seed_system.clear();
  for (size_t g = 0; g < G; ++g) {
    for (size_t j = 0; j < P + 2; ++j) {
      const rng_state& s = all[g * (P + 2) + j];
      auto& dst = j < 2 ? seed_filter : seed_system;
      dst.insert(dst.end(), s.s, s.s + 4);
    }
  }
}

template <typename T>
int filter_run_impl(const double* pars, size_t G, size_t P, double time_start,
                    const double* time, size_t n_time, double dt,
                    const double* data, size_t n_groups_data,
                    const uint64_t* rng_in, uint64_t seed, int n_threads,
                    const double* initial, const size_t* index_group,
                    size_t n_index, double* ll_out, double* state_out,
                    uint64_t* rng_out, uint32_t* ancestry, double* traj, char* err,
                    size_t errlen) {
  try {
    const size_t n_streams = G * (P + 2);
    std::vector<uint64_t> seed_all;
    if (rng_in) {
      seed_all.assign(rng_in, rng_in + 4 * n_streams);
    } else {
      seed_all.resize(4);
      seed_words(seed, seed_all.data());
    }
    const auto streams = prng(n_streams, seed_all, false);
    std::vector<uint64_t> seed_filter, seed_system;
    split_filter_streams(streams, G, P, seed_filter, seed_system);
    discrete_system<T> sys(read_shared<T>(pars, G), time_start, dt, P,
                           seed_system, false, n_threads);
    filter<T> obj(sys, time_start, std::vector<double>(time, time + n_time),
                  read_data<T>(data, n_time * n_groups_data), seed_filter);
    if (initial) {
      obj.sys.set_state(initial);
    }
    const auto idx = make_index(index_group, n_index, G);
    obj.run(initial == nullptr, idx, ancestry, traj);
    for (size_t i = 0; i < G; ++i) ll_out[i] = obj.log_likelihood()[i];
    if (state_out) {
      std::copy(obj.sys.state().begin(), obj.sys.state().end(), state_out);
    }
    if (rng_out) {
      for (size_t g = 0; g < G; ++g) {
        for (size_t j = 0; j < P + 2; ++j) {
          const rng_state& s = j < 2 ? obj.rng()[2 * g + j]
                                     : obj.sys.rng()[g * P + (j - 2)];
          std::memcpy(rng_out + 4 * (g * (P + 2) + j), s.s, 32);
        }
      }
    }
    return 0;
  } catch (const std::exception& e) {
    return fail(e, err, errlen);
  }
}

template <typename T>
int simulate_impl(const double* pars, size_t G, size_t P, double time_start,
                  double dt, const double* times, size_t n_times,
                  const uint64_t* rng_in, int deterministic, int n_threads,
                  double* out, uint64_t* rng_out, char* err, size_t errlen) {
  try {
    std::vector<uint64_t> seed(rng_in, rng_in + 4 * G * P);
    discrete_system<T> sys(read_shared<T>(pars, G), time_start, dt, P, seed,
                           deterministic != 0, n_threads);
    sys.set_state_initial();
    std::vector<size_t> all(G);
    std::iota(all.begin(), all.end(), 0);
    const size_t stride = G * P * T::n_state;
    for (size_t t = 0; t < n_times; ++t) {
      sys.run_to_time(times[t], all);
      std::copy(sys.state().begin(), sys.state().end(), out + t * stride);
    }
    if (rng_out) {
      for (size_t k = 0; k < G * P; ++k) {
        std::memcpy(rng_out + 4 * k, sys.rng()[k].s, 32);
      }
    }
    return 0;
  } catch (const std::exception& e) {
    return fail(e, err, errlen);
  }
}

// Deterministic likelihood of a discrete-time model: unfilter<T>::run
// (dust2/unfilter.hpp:42-72) over a system whose streams are deterministic
// (dust2/r/discrete/unfilter.hpp:38-46): per data time run_to_time,
// compare_data and ll[k] += ll_step[k] for every particle, no resampling.
// data: n_time x G_data x n_data; ll_out: G*P; state_out: G x P x S or NULL;
// traj_out: n_time x G x P x S or NULL.
template <typename T>
int discrete_unfilter_impl(const double* pars, size_t G, size_t P, double time_start,
                           const double* time, size_t n_time, double dt, const double* data,
                           size_t n_groups_data, int n_threads, double* ll_out,
                           double* state_out, double* traj_out, char* err, size_t errlen) {
  try {
    std::vector<uint64_t> seed(4);
    seed_words(42, seed.data());
    discrete_system<T> sys(read_shared<T>(pars, G), time_start, dt, P, seed, true,
                           n_threads);
    const auto d = read_data<T>(data, n_time * n_groups_data);
    std::vector<size_t> all(G);
    std::iota(all.begin(), all.end(), 0);
    sys.set_state_initial();
    std::vector<double> ll(G * P, 0.0), ll_step(G * P);
    const size_t stride = G * P * T::n_state;
    for (size_t t = 0; t < n_time; ++t) {
      sys.run_to_time(time[t], all);
      sys.compare_data(d.data() + t * n_groups_data, n_groups_data, all, ll_step.data());
      for (size_t k = 0; k < ll.size(); ++k) ll[k] += ll_step[k];
      if (traj_out) std::copy(sys.state().begin(), sys.state().end(), traj_out + t * stride);
    }
    std::copy(ll.begin(), ll.end(), ll_out);
    if (state_out) std::copy(sys.state().begin(), sys.state().end(), state_out);
    return 0;
  } catch (const std::exception& e) {
    return fail(e, err, errlen);
  }
}

}  // namespace

extern "C" {

// ---- RNG ------------------------------------------------------------------
void oracle_seed_words(uint64_t seed, uint64_t* out4) { seed_words(seed, out4); }

void oracle_prng(const uint64_t* seed, size_t n_seed, size_t n, uint64_t* out) {
  std::vector<uint64_t> s(seed, seed + 4 * n_seed);
  const auto streams = prng(n, s, false);
  for (size_t i = 0; i < n; ++i) std::memcpy(out + 4 * i, streams[i].s, 32);
}

void oracle_jump(uint64_t* s4, int long_jump) {
  rng_state r{{s4[0], s4[1], s4[2], s4[3]}, false};
  if (long_jump) {
    rng_long_jump(r);
  } else {
    rng_jump(r);
  }
  std::memcpy(s4, r.s, 32);
}

void oracle_rng_next(uint64_t* s4, size_t n, uint64_t* out) {
  rng_state r{{s4[0], s4[1], s4[2], s4[3]}, false};
  for (size_t i = 0; i < n; ++i) out[i] = rng_next(r);
  std::memcpy(s4, r.s, 32);
}

void oracle_random_real(uint64_t* s4, size_t n, double* out) {
  rng_state r{{s4[0], s4[1], s4[2], s4[3]}, false};
  for (size_t i = 0; i < n; ++i) out[i] = random_real(r);
  std::memcpy(s4, r.s, 32);
}

// kind: 0 binomial(a=n, b=p), 1 exponential_rate(a=rate), 2 normal(a=mean,
// b=sd), 3 poisson(a=lambda).  Draws n_draws values from ONE stream,
// with per-draw arguments.
int oracle_sample(int kind, uint64_t* s4, int deterministic, size_t n_draws,
                  const double* a, const double* b, double* out, char* err,
                  size_t errlen) {
  rng_state r{{s4[0], s4[1], s4[2], s4[3]}, deterministic != 0};
  try {
    for (size_t i = 0; i < n_draws; ++i) {
      switch (kind) {
        case 0: out[i] = binomial(r, a[i], b[i]); break;
        case 1: out[i] = exponential_rate(r, a[i]); break;
        case 2: out[i] = normal(r, a[i], b[i]); break;
        default: out[i] = poisson(r, a[i]); break;
      }
    }
  } catch (const std::exception& e) {
    std::memcpy(s4, r.s, 32);
    return fail(e, err, errlen);
  }
  std::memcpy(s4, r.s, 32);
  return 0;
}

double oracle_dpois_log(double x, double lambda) { return dpois_log(x, lambda); }

double oracle_scale_log_weights(size_t n, double* w) {
  return scale_log_weights(n, w);
}

void oracle_resample_weight(size_t n, const double* w, double u, uint32_t* idx) {
  resample_weight(n, w, u, idx);
}

// ---- filters --------------------------------------------------------------
// pars: G x n_pars (sir: N, I0, beta, gamma, exp_noise); data: n_time x
// n_groups_data x n_data; rng_in: G*(P+2) streams (NULL: expand `seed`);
// state_out: AoS [(g*P + p)*S + i]; rng_out: G*(P+2) streams, filter
// interleaved; ancestry: n_time x G x P (or NULL).
int oracle_sir_filter_run(const double* pars, size_t G, size_t P,
                          double time_start, const double* time, size_t n_time,
                          double dt, const double* data, size_t n_groups_data,
                          const uint64_t* rng_in, uint64_t seed, int n_threads,
                          const double* initial, const size_t* index_group,
                          size_t n_index, double* ll_out, double* state_out,
                          uint64_t* rng_out, uint32_t* ancestry, double* traj,
                          char* err, size_t errlen) {
  return filter_run_impl<sir>(pars, G, P, time_start, time, n_time, dt, data,
                              n_groups_data, rng_in, seed, n_threads, initial,
                              index_group, n_index, ll_out, state_out, rng_out,
                              ancestry, traj, err, errlen);
}

int oracle_sir_tv_filter_run(const double* pars, size_t G, size_t P, double time_start,
                             const double* time, size_t n_time, double dt, const double* data,
                             size_t n_groups_data, const uint64_t* rng_in, uint64_t seed,
                             int n_threads, const double* initial, const size_t* index_group,
                             size_t n_index, double* ll_out, double* state_out,
                             uint64_t* rng_out, uint32_t* ancestry, double* traj, char* err,
                             size_t errlen) {
  return filter_run_impl<sir_tv>(pars, G, P, time_start, time, n_time, dt, data, n_groups_data,
                                 rng_in, seed, n_threads, initial, index_group, n_index, ll_out,
                                 state_out, rng_out, ancestry, traj, err, errlen);
}

int oracle_seir16_filter_run(const double* pars, size_t G, size_t P,
                             double time_start, const double* time,
                             size_t n_time, double dt, const double* data,
                             size_t n_groups_data, const uint64_t* rng_in,
                             uint64_t seed, int n_threads, const double* initial,
                             const size_t* index_group, size_t n_index,
                             double* ll_out, double* state_out, uint64_t* rng_out,
                             uint32_t* ancestry, double* traj, char* err, size_t errlen) {
  return filter_run_impl<seir16>(pars, G, P, time_start, time, n_time, dt, data,
                                 n_groups_data, rng_in, seed, n_threads, initial,
                                 index_group, n_index, ll_out, state_out,
                                 rng_out, ancestry, traj, err, errlen);
}

// System simulate (dust2/discrete/system.hpp:121-128): state at each time,
// out: n_times x G x P x S.  rng_in: G*P streams.
int oracle_sir_simulate(const double* pars, size_t G, size_t P,
                        double time_start, double dt, const double* times,
                        size_t n_times, const uint64_t* rng_in,
                        int deterministic, int n_threads, double* out,
                        uint64_t* rng_out, char* err, size_t errlen) {
  return simulate_impl<sir>(pars, G, P, time_start, dt, times, n_times, rng_in,
                            deterministic, n_threads, out, rng_out, err, errlen);
}

int oracle_sir_tv_simulate(const double* pars, size_t G, size_t P, double time_start,
                           double dt, const double* times, size_t n_times,
                           const uint64_t* rng_in, int deterministic, int n_threads,
                           double* out, uint64_t* rng_out, char* err, size_t errlen) {
  return simulate_impl<sir_tv>(pars, G, P, time_start, dt, times, n_times, rng_in,
                               deterministic, n_threads, out, rng_out, err, errlen);
}

int oracle_seir16_simulate(const double* pars, size_t G, size_t P,
                           double time_start, double dt, const double* times,
                           size_t n_times, const uint64_t* rng_in,
                           int deterministic, int n_threads, double* out,
                           uint64_t* rng_out, char* err, size_t errlen) {
  return simulate_impl<seir16>(pars, G, P, time_start, dt, times, n_times,
                               rng_in, deterministic, n_threads, out, rng_out,
                               err, errlen);
}

// Deterministic unfilter of a discrete-time model (discrete_unfilter_impl).
int oracle_sir_unfilter_det_run(const double* pars, size_t G, size_t P, double time_start,
                                const double* time, size_t n_time, double dt,
                                const double* data, size_t n_groups_data, int n_threads,
                                double* ll_out, double* state_out, double* traj_out, char* err,
                                size_t errlen) {
  return discrete_unfilter_impl<sir>(pars, G, P, time_start, time, n_time, dt, data,
                                     n_groups_data, n_threads, ll_out, state_out, traj_out, err,
                                     errlen);
}

int oracle_seir16_unfilter_run(const double* pars, size_t G, size_t P, double time_start,
                               const double* time, size_t n_time, double dt,
                               const double* data, size_t n_groups_data, int n_threads,
                               double* ll_out, double* state_out, double* traj_out, char* err,
                               size_t errlen) {
  return discrete_unfilter_impl<seir16>(pars, G, P, time_start, time, n_time, dt, data,
                                        n_groups_data, n_threads, ll_out, state_out, traj_out,
                                        err, errlen);
}

// The random walk as a system (walk.cpp), len walkers in 1..8.
int oracle_walk_simulate(size_t len, const double* pars, size_t G, size_t P, double time_start,
                         double dt, const double* times, size_t n_times, const uint64_t* rng_in,
                         int deterministic, int n_threads, double* out, uint64_t* rng_out,
                         char* err, size_t errlen) {
  switch (len) {
#define D2_WALK(L)                                                                         \
  case L:                                                                                  \
    return simulate_impl<walk<L>>(pars, G, P, time_start, dt, times, n_times, rng_in,      \
                                  deterministic, n_threads, out, rng_out, err, errlen);
    D2_WALK(1)
    D2_WALK(2)
    D2_WALK(3)
    D2_WALK(4)
    D2_WALK(5)
    D2_WALK(6)
    D2_WALK(7)
    D2_WALK(8)
#undef D2_WALK
    default:
      if (err && errlen) std::snprintf(err, errlen, "walk: len must be in 1..8");
      return 1;
  }
}

// ---- ODE unfilter ---------------------------------------------------------
// control: [max_steps, atol, rtol, step_size_min, step_size_max].
// stats_out (G*P x 4, or NULL): n_steps, n_steps_accepted,
// n_steps_rejected, step_size after the run.
int oracle_sirode_unfilter_run(const double* pars, size_t G, size_t P,
                               double time_start, const double* time,
                               size_t n_time, const double* data,
                               size_t n_groups_data, const double* control,
                               int n_threads, double* ll_out, double* state_out,
                               double* stats_out, char* err, size_t errlen) {
  try {
    ode_control ctl;
    ctl.max_steps = static_cast<size_t>(control[0]);
    ctl.atol = control[1];
    ctl.rtol = control[2];
    ctl.step_size_min = std::max(control[3], std::numeric_limits<double>::epsilon());
    ctl.step_size_max = control[4];
    continuous_system<sirode> sys(read_shared<sirode>(pars, G), time_start, P, ctl,
                                  n_threads);
    const auto d = read_data<sirode>(data, n_time * n_groups_data);
    unfilter_run(sys, time_start, std::vector<double>(time, time + n_time),
                 d.data(), n_groups_data, ll_out);
    if (state_out) {
      std::copy(sys.state().begin(), sys.state().end(), state_out);
    }
    if (stats_out) {
      for (size_t k = 0; k < G * P; ++k) {
        const auto& in = sys.ode_internals_()[k];
        stats_out[4 * k + 0] = in.n_steps;
        stats_out[4 * k + 1] = in.n_steps_accepted;
        stats_out[4 * k + 2] = in.n_steps_rejected;
        stats_out[4 * k + 3] = in.step_size;
      }
    }
    return 0;
  } catch (const std::exception& e) {
    return fail(e, err, errlen);
  }
}

// ---- deterministic discrete SIR unfilter + adjoint gradient ---------------
// (dust2/unfilter.hpp:77-229, dust2/discrete/system.hpp:290-482).  P = 1;
// data per group (n_time x G).  grad_out: G x 3 (beta, gamma, I0) or NULL.
int oracle_sir_unfilter_run(const double* pars, size_t G, double time_start,
                            const double* time, size_t n_time, double dt,
                            const double* data, double* ll_out,
                            double* grad_out, int exact, char* err, size_t errlen) {
  try {
    constexpr size_t S = sir::n_state;
    constexpr size_t A = S + sir::n_gradient;
    std::vector<uint64_t> seed(4);
    seed_words(42, seed.data());
    discrete_system<sir> sys(read_shared<sir>(pars, G), time_start, dt, 1, seed,
                             true, 1);
    const auto d = read_data<sir>(data, n_time * G);
    std::vector<size_t> offset(n_time + 1, 0);
    for (size_t i = 0; i < n_time; ++i) {
      offset[i + 1] = std::round(std::max(0.0, time[i] - time_start) / dt) * G * S;
    }
    std::vector<double> history(offset.back() + G * S, 0.0);
    sys.set_time(time_start);
    sys.set_state_initial();
    sys.run_to_time_history(time[n_time - 1], history.data());
    std::fill(ll_out, ll_out + G, 0.0);
    rng_state det{{0, 0, 0, 0}, true};
    for (size_t i = 0; i < n_time; ++i) {
      const double* st = history.data() + offset[i + 1];
      for (size_t g = 0; g < G; ++g) {
        ll_out[g] += sir::compare_data(time[i], st + g * S, d[i * G + g],
                                       sys.shared()[g], det);
      }
    }
    if (grad_out) {
      std::vector<double> curr(A * G, 0.0), next(A * G, 0.0);
      double* ac = curr.data();
      double* an = next.data();
      const size_t stride = G * S;
      double tcur = time[n_time - 1];
      for (size_t irev = 0; irev < n_time; ++irev) {
        const size_t i = n_time - irev - 1;
        const double time_i = i == 0 ? time_start : time[i - 1];
        const double* state_i = history.data() + offset[i + 1];
        for (size_t g = 0; g < G; ++g) {
          sir::adjoint_compare_data(state_i + g * S, ac + g * A, d[i * G + g],
                                    sys.shared()[g], an + g * A, exact != 0);
        }
        std::swap(ac, an);
        const size_t n_steps = std::round(std::max(0.0, tcur - time_i) / dt);
        for (size_t g = 0; g < G; ++g) {
          double* c = ac + g * A;
          double* n = an + g * A;
          for (size_t s = 1; s <= n_steps; ++s) {
            const double time_s = tcur - s * dt;
            const double* st = state_i + g * S - s * stride;
            sir::adjoint_update(dt, st, c, sys.shared()[g], n, exact != 0);
            std::swap(c, n);
            if (std::fmod(time_s, sir::zero_period) == 0) {
              c[sir::zero_first] = 0;
            }
          }
        }
        if (n_steps % 2 == 1) {
          std::swap(ac, an);
        }
        tcur = time_i;
      }
      for (size_t g = 0; g < G; ++g) {
        sir::adjoint_initial(history.data() + g * S, ac + g * A, sys.shared()[g],
                             an + g * A, exact != 0);
      }
      for (size_t g = 0; g < G; ++g) {
        for (size_t j = 0; j < sir::n_gradient; ++j) {
          grad_out[g * sir::n_gradient + j] = an[g * A + S + j];
        }
      }
    }
    return 0;
  } catch (const std::exception& e) {
    return fail(e, err, errlen);
  }
}

// Trajectory export with reorder (dust2/trajectories.hpp:128-189, using
// reorder_state_group_ / reorder_state_single_ at :313-343): walking back
// from the last time, write the state of each tracked particle, then
// replace the particle by its ancestor at that time.  raw [T][G][P][S],
// ancestry [T][G][P] (NULL: never reordered), select [G] (NULL: every
// particle); out [T][G][P or 1][S].
void oracle_trajectory_lineage(const double* raw, const uint32_t* ancestry, size_t T,
                               size_t G, size_t P, size_t S, const size_t* select,
                               double* out) {
  const size_t p_out = select ? 1 : P;
  std::vector<size_t> index(G * p_out);
  for (size_t g = 0; g < G; ++g) {
    for (size_t k = 0; k < p_out; ++k) index[g * p_out + k] = select ? select[g] : k;
  }
  for (size_t irev = 0; irev < T; ++irev) {
    const size_t t = T - irev - 1;
    for (size_t g = 0; g < G; ++g) {
      for (size_t k = 0; k < p_out; ++k) {
        size_t& i = index[g * p_out + k];
        std::copy_n(raw + ((t * G + g) * P + i) * S, S, out + ((t * G + g) * p_out + k) * S);
        if (ancestry) i = ancestry[(t * G + g) * P + i];
      }
    }
  }
}

// Log-likelihood of the SIR ODE unfilter and its gradient with respect to
// (beta, gamma, I0) by forward sensitivities (sirode_sens): one augmented
// integration per parameter; the likelihood is that of the plain unfilter.
// d/dlambda log Poisson(x; lambda) = x / lambda - 1.
int oracle_sirode_unfilter_gradient(const double* pars, size_t G, double time_start,
                                    const double* time, size_t n_time, const double* data,
                                    size_t n_groups_data, const double* control_in,
                                    int n_threads, double* ll_out, double* grad_out,
                                    char* err, size_t errlen) {
  try {
    ode_control ctl;
    if (control_in) {
      ctl.max_steps = static_cast<size_t>(control_in[0]);
      ctl.atol = control_in[1];
      ctl.rtol = control_in[2];
      ctl.step_size_min = std::max(control_in[3], std::numeric_limits<double>::epsilon());
      ctl.step_size_max = control_in[4];
    }
    std::string failure;
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(n_threads) collapse(2)
#endif
    for (size_t g = 0; g < G; ++g) {
      for (int param = 0; param < 3; ++param) {
        try {
          sirode_sens::shared_state sh{sirode::build_shared(pars + g * sirode::n_pars), param};
          double y[sirode_sens::n_state];
          sirode_sens::initial(sh, y);
          ode_internals<sirode_sens::n_state> in;
          ode_solver<sirode_sens> solver(ctl);
          solver.initialise(time_start, y, in, sh);
          double t = time_start;
          double ll = 0, dll = 0;
          for (size_t i = 0; i < n_time; ++i) {
            solver.run(t, time[i], y, in, sh);
            t = time[i];
            const double x = data[i * n_groups_data + (n_groups_data == 1 ? 0 : g)];
            if (std::isnan(x)) continue;
            const double lambda = y[8];
            ll += dpois_log(x, lambda);
            dll += (x == 0 ? -1.0 : x / lambda - 1) * y[9];
          }
          if (param == 0) ll_out[g] = ll;
          grad_out[3 * g + param] = dll;
        } catch (const std::exception& e) {
#ifdef _OPENMP
#pragma omp critical
#endif
          failure = e.what();
        }
      }
    }
    if (!failure.empty()) throw std::runtime_error(failure);
    return 0;
  } catch (const std::exception& e) {
    return fail(e, err, errlen);
  }
}

// Interpolation (oracle::interp): values out[n_z * n_series]; returns 0, or
// 1 / 2 for the first time z that lies before / after the knots (its index
// in *bad).
int oracle_interpolate(int kind, const double* t, size_t n_t, const double* y,
                       size_t n_series, const double* z, size_t n_z, double* out,
                       size_t* bad) {
  const std::vector<double> tv(t, t + n_t), yv(y, y + n_t * n_series);
  const std::vector<double> k = kind == oracle::interp::spline
                                    ? oracle::interp::spline_slopes(tv, yv, n_series)
                                    : std::vector<double>(yv.size(), 0.0);
  for (size_t q = 0; q < n_z; ++q) {
    const int rc = oracle::interp::eval(kind, tv, yv, k, n_series, z[q], out + q * n_series);
    if (rc != 0) {
      *bad = q;
      return rc;
    }
  }
  return 0;
}

}  // extern "C"
