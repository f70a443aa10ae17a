// ORACLE — TEST INFRASTRUCTURE ONLY.
//
// A plain C++17/OpenMP FP64 CPU restatement of dust2's batched particle
// engine. Only the tests, the smoke check and the CPU-baseline leg of the
// benchmark may load it; the product (dust2_amd/) never links, calls or
// falls back to it.
//
// It mirrors the reference's *structure* on purpose, because it is also the
// CPU baseline: AoS particle state [(P*g + p)*S + i], double buffered
// (dust2/discrete/system.hpp:45-46), OpenMP `parallel for schedule(static)
// collapse(2)` over group x particle (dust2/discrete/system.hpp:67-68),
// serial per-group log-sum-exp and systematic resampling
// (dust2/filter.hpp:141-160, dust2/filter_details.hpp:11-74) and a serial
// gather reorder (dust2/discrete/system.hpp:169-181).
//
// Pinning status of each part (what checks it independently of this file):
//   xoshiro256+ next/jump/long_jump ... pinned: jump == T^(2^128) and
//       long_jump == T^(2^192) by GF(2) matrix squaring; splitmix64 known
//       answers (tests/test_oracle_rng.py).
//   seed_words (integer seed -> 4 words) ... parity unpinned (monty's
//       as_rng_seed is not in the reference tree; structure pinned only).
//   random_real ... parity unpinned (monty); range/resolution checked.
//   binomial / exponential / normal / poisson draws ... parity unpinned
//       (monty); distributions pinned by exact-pmf chi-square and KS tests.
//   dpois_log ... pinned against scipy.stats.poisson.logpmf.
//   scale_log_weights / resample_weight ... pinned against the reference's
//       own filter_details.hpp compiled into oracle/_ref, and closed forms.
//   dp5 solver ... pinned against the reference's own continuous/solver.hpp
//       compiled into oracle/_ref (daily states, step counters, step size),
//       and the logistic closed form.
//   filter / unfilter orchestration ... pinned by the reference's invariant
//       tests restated in tests/ (grouped == independent, NaN data, P == 1,
//       deterministic closed form).
//   seir16 model ... no reference model exists: invariants only.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <limits>
#include <numeric>
#include <sstream>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace oracle {

// ---------------------------------------------------------------------------
// xoshiro256+ (Blackman & Vigna), the generator behind
// monty::random::generator<double> used by every dust2 model
// (inst/examples/sir.cpp:32; 32 bytes per stream, tests/testthat/test-sir.R:14)
// ---------------------------------------------------------------------------
struct rng_state {
  uint64_t s[4];
  bool deterministic;
};

inline uint64_t rotl64(uint64_t x, int k) {
  return (x << k) | (x >> (64 - k));
}

inline uint64_t rng_next(rng_state& r) {
  const uint64_t out = r.s[0] + r.s[3];
  const uint64_t t = r.s[1] << 17;
  r.s[2] ^= r.s[0];
  r.s[3] ^= r.s[1];
  r.s[1] ^= r.s[2];
  r.s[0] ^= r.s[3];
  r.s[2] ^= t;
  r.s[3] = rotl64(r.s[3], 45);
  return out;
}

// Published jump polynomials: 2^128 and 2^192 steps ahead.
constexpr uint64_t kJump[4] = {0x180ec6d33cfd0abaULL, 0xd5a61266f0c9392cULL,
                               0xa9582618e03fc9aaULL, 0x39abdc4529b1661cULL};
constexpr uint64_t kLongJump[4] = {0x76e15d3efefdcbbfULL, 0xc5004e441c522fb3ULL,
                                   0x77710069854ee241ULL, 0x39109bb02acbe635ULL};

inline void rng_apply_poly(rng_state& r, const uint64_t* poly) {
  uint64_t acc[4] = {0, 0, 0, 0};
  for (int w = 0; w < 4; ++w) {
    for (int b = 0; b < 64; ++b) {
      if (poly[w] & (uint64_t(1) << b)) {
        for (int i = 0; i < 4; ++i) acc[i] ^= r.s[i];
      }
      rng_next(r);
    }
  }
  std::memcpy(r.s, acc, sizeof(acc));
}

inline void rng_jump(rng_state& r) { rng_apply_poly(r, kJump); }
inline void rng_long_jump(rng_state& r) { rng_apply_poly(r, kLongJump); }

// Integer seed -> first stream (vignettes/design.Rmd:72-75: "we seed the
// first stream using the splitmix64 RNG").  Each word is the splitmix64
// mix of the previous word plus the golden gamma.
inline uint64_t splitmix64_mix(uint64_t x) {
  uint64_t z = x + 0x9e3779b97f4a7c15ULL;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

inline void seed_words(uint64_t seed, uint64_t* out4) {
  for (int i = 0; i < 4; ++i) {
    seed = splitmix64_mix(seed);
    out4[i] = seed;
  }
}

// prng(n, seed, deterministic): streams present in `seed` are taken
// verbatim; each further stream is the previous one jumped once
// (vignettes/design.Rmd:75; dust2/r/discrete/filter.hpp:63-64 expands
// G*(P+2) streams this way).
inline std::vector<rng_state> prng(size_t n, const std::vector<uint64_t>& seed,
                                   bool deterministic) {
  std::vector<rng_state> ret(n);
  const size_t n_seed = seed.size() / 4;
  rng_state s{{0, 0, 0, 0}, deterministic};
  for (size_t i = 0; i < n; ++i) {
    if (i < n_seed) {
      std::memcpy(s.s, seed.data() + 4 * i, 32);
    } else {
      rng_jump(s);
    }
    s.deterministic = deterministic;
    ret[i] = s;
  }
  return ret;
}

// U[0,1) on the 2^-53 grid from the top 53 bits of one draw.
inline double random_real(rng_state& r) {
  return static_cast<double>(rng_next(r) >> 11) * 0x1.0p-53;
}

// ---------------------------------------------------------------------------
// Samplers (monty::random::*; external to the reference tree).  Written from
// the published algorithms; the draw sequences are documented choices.
// ---------------------------------------------------------------------------
inline std::string binomial_error(double n, double p) {
  char buf[256];
  std::snprintf(buf, sizeof(buf),
                "Invalid call to binomial with n = %.0f, p = %g, q = %g", n, p,
                1 - p);
  return buf;
}

// Remainder of Stirling's series: log(k!) - [(k+1/2)log(k+1) - (k+1) +
// log(2 pi)/2]; exact values (to double rounding) for k <= 9, the
// asymptotic series 1/12 - 1/360 x^-2 + 1/1260 x^-4 over x = k+1 above.
constexpr double kStirlingTail[10] = {
    0.08106146679532726,  0.0413406959554093,   0.02767792568499834,
    0.020790672103765093, 0.016644691189821193, 0.013876128823070748,
    0.01189670994589177,  0.010411265261972096, 0.009255462182712733,
    0.00833056343336287};

inline double stirling_tail(double k) {
  if (k <= 9) {
    return kStirlingTail[static_cast<int>(k)];
  }
  const double x = k + 1;
  const double x2 = x * x;
  return (1.0 / 12 - (1.0 / 360 - 1.0 / 1260 / x2) / x2) / x;
}

// Inversion by sequential search of the pmf, for n*p < 10 (p <= 1/2):
// pmf(k) = pmf(k-1) * (p/q) * (n+1-k)/k.  Restarts with a fresh uniform if
// rounding lets the search run past n.
inline double binomial_inversion(rng_state& rng, double n, double p) {
  const double q = 1 - p;
  const double r = p / q;
  const double g = r * (n + 1);
  while (true) {
    double u = random_real(rng);
    double f = std::pow(q, n);
    double k = 0;
    bool ok = true;
    while (u >= f) {
      u -= f;
      k += 1;
      if (k > n) {
        ok = false;
        break;
      }
      f *= (g / k - r);
    }
    if (ok) {
      return k;
    }
  }
}

// BTRS: transformed rejection with squeeze, Hormann (1993) "The generation
// of binomial random variates", J. Stat. Comput. Simul. 46:101-110.
// p <= 1/2 and n*p >= 10.  One uniform for the transform (u), one for the
// acceptance (v), in that order.
inline double binomial_btrs(rng_state& rng, double n, double p) {
  const double q = 1 - p;
  const double spq = std::sqrt(n * p * q);
  const double b = 1.15 + 2.53 * spq;
  const double a = -0.0873 + 0.0248 * b + 0.01 * p;
  const double c = n * p + 0.5;
  const double v_r = 0.92 - 4.2 / b;
  const double alpha = (2.83 + 5.1 / b) * spq;
  const double lpq = std::log(p / q);
  const double m = std::floor((n + 1) * p);
  const double h = (m + 0.5) * std::log(m + 1) + stirling_tail(m) +
                   (n - m + 0.5) * std::log(n - m + 1) + stirling_tail(n - m);
  while (true) {
    const double u = random_real(rng) - 0.5;
    double v = random_real(rng);
    const double us = 0.5 - std::fabs(u);
    const double k = std::floor((2 * a / us + b) * u + c);
    if (us >= 0.07 && v <= v_r) {
      return k;
    }
    if (k < 0 || k > n) {
      continue;
    }
    v = std::log(v * alpha / (a / (us * us) + b));
    const double bound = h - (k + 0.5) * std::log(k + 1) - stirling_tail(k) -
                         (n - k + 0.5) * std::log(n - k + 1) -
                         stirling_tail(n - k) + (k - m) * lpq;
    if (v <= bound) {
      return k;
    }
  }
}

// Binomial(n, p); deterministic streams return the mean n*p.  Invalid
// arguments throw with the reference's message (tests/testthat/test-sir.R:259).
inline double binomial(rng_state& rng, double n, double p) {
  if (n < 0 || p < 0 || p > 1) {
    throw std::runtime_error(binomial_error(n, p));
  }
  if (rng.deterministic) {
    return n * p;
  }
  if (n == 0 || p == 0) {
    return 0;
  }
  if (p == 1) {
    return n;
  }
  const bool flip = p > 0.5;
  const double q = flip ? 1 - p : p;
  const double k = (n * q >= 10) ? binomial_btrs(rng, n, q)
                                 : binomial_inversion(rng, n, q);
  return flip ? n - k : k;
}

// Exp(rate) by inversion: -log(1 - U)/rate, 1 - U exact and in (0, 1].
// Deterministic streams return the mean 1/rate
// (tests/testthat/test-unfilter.R:22 relies on the +1e-6 this gives).
inline double exponential_rate(rng_state& rng, double rate) {
  if (rng.deterministic) {
    return 1 / rate;
  }
  return -std::log(1 - random_real(rng)) / rate;
}

// Standard normal by the Box-Muller transform (used by the random-walk
// model family, inst/examples/walk.cpp:57).
inline double normal(rng_state& rng, double mean, double sd) {
  if (rng.deterministic) {
    return mean;
  }
  double u1;
  double u2;
  do {
    u1 = random_real(rng);
    u2 = random_real(rng);
  } while (u1 <= std::numeric_limits<double>::min());
  const double z = std::sqrt(-2 * std::log(u1)) * std::cos(2 * M_PI * u2);
  return mean + sd * z;
}

// Poisson(lambda): multiplication method below 10, PTRS (Hormann 1993,
// "The transformed rejection method for generating Poisson random
// variables", Insurance: Mathematics and Economics 12:39-45) above.
inline double poisson(rng_state& rng, double lambda) {
  if (rng.deterministic) {
    return lambda;
  }
  if (lambda == 0) {
    return 0;
  }
  if (lambda < 10) {
    const double l = std::exp(-lambda);
    double k = 0;
    double prod = random_real(rng);
    while (prod > l) {
      k += 1;
      prod *= random_real(rng);
    }
    return k;
  }
  const double slam = std::sqrt(lambda);
  const double loglam = std::log(lambda);
  const double b = 0.931 + 2.53 * slam;
  const double a = -0.059 + 0.02483 * b;
  const double inv_alpha = 1.1239 + 1.1328 / (b - 3.4);
  const double v_r = 0.9277 - 3.6224 / (b - 2);
  while (true) {
    const double u = random_real(rng) - 0.5;
    const double v = random_real(rng);
    const double us = 0.5 - std::fabs(u);
    const double k = std::floor((2 * a / us + b) * u + lambda + 0.43);
    if (us >= 0.07 && v <= v_r) {
      return k;
    }
    if (k < 0 || (us < 0.013 && v > us)) {
      continue;
    }
    if (std::log(v) + std::log(inv_alpha) - std::log(a / (us * us) + b) <=
        -lambda + k * loglam - std::lgamma(k + 1)) {
      return k;
    }
  }
}

// log Poisson pmf (monty::density::poisson(x, lambda, true),
// inst/examples/sir.cpp:119); equals R's dpois(x, lambda, log = TRUE).
inline double dpois_log(double x, double lambda) {
  if (x == 0 && lambda == 0) {
    return 0;
  }
  if (lambda == 0) {
    return -std::numeric_limits<double>::infinity();
  }
  return x * std::log(lambda) - lambda - std::lgamma(x + 1);
}

// ---------------------------------------------------------------------------
// Per-particle error capture (dust2/errors.hpp:10-60)
// ---------------------------------------------------------------------------
class errors {
 public:
  explicit errors(size_t n = 0) : err_(n), seen_(n, 0), count_(0) {}
  void reset() {
    count_ = 0;
    std::fill(seen_.begin(), seen_.end(), 0);
    std::fill(err_.begin(), err_.end(), std::string());
  }
  bool unresolved() const { return count_ > 0; }
  void capture(const std::exception& e, size_t i) {
    err_[i] = e.what();
    seen_[i] = 1;
  }
  void report(const char* title = "particles", size_t n_max = 4) {
    count_ = std::accumulate(seen_.begin(), seen_.end(), size_t(0));
    if (count_ == 0) {
      return;
    }
    std::stringstream msg;
    msg << count_ << " " << title << " reported errors.";
    const size_t n_report = std::min(n_max, count_);
    for (size_t i = 0, j = 0; i < seen_.size() && j < n_report; ++i) {
      if (seen_[i]) {
        msg << std::endl << "  - " << i + 1 << ": " << err_[i];
        ++j;
      }
    }
    if (n_report < count_) {
      msg << std::endl << "  - (and " << (count_ - n_report) << " more)";
    }
    throw std::runtime_error(msg.str());
  }

 private:
  std::vector<std::string> err_;
  std::vector<char> seen_;
  size_t count_;
};

// ---------------------------------------------------------------------------
// Models, written to dust2's generated-class contract
// ---------------------------------------------------------------------------

// inst/examples/sir.cpp:12-199
struct sir {
  static constexpr size_t n_state = 5;
  static constexpr size_t n_pars = 5;
  struct shared_state {
    double N, I0, beta, gamma, exp_noise;
  };
  struct data_type {
    double incidence;
  };
  static shared_state build_shared(const double* p) {
    return shared_state{p[0], p[1], p[2], p[3], p[4]};
  }
  static constexpr size_t n_data = 1;
  static data_type build_data(const double* d) { return data_type{d[0]}; }
  // zero_every = {{1, {4}}} (sir.cpp:95-97)
  static constexpr double zero_period = 1;
  static constexpr size_t zero_first = 4;
  static constexpr size_t zero_count = 1;

  static void initial(double, const shared_state& sh, rng_state&, double* y) {
    y[0] = sh.N - sh.I0;
    y[1] = sh.I0;
    y[2] = 0;
    y[3] = 0;
    y[4] = 0;
  }
  static void update(double, double dt, const double* y, const shared_state& sh,
                     rng_state& rng, double* yn) {
    const double S = y[0];
    const double I = y[1];
    const double R = y[2];
    const double cases_cumul = y[3];
    const double cases_inc = y[4];
    const double p_SI = 1 - std::exp(-sh.beta * I / sh.N * dt);
    const double p_IR = 1 - std::exp(-sh.gamma * dt);
    const double n_SI = binomial(rng, S, p_SI);
    const double n_IR = binomial(rng, I, p_IR);
    yn[0] = S - n_SI;
    yn[1] = I + n_SI - n_IR;
    yn[2] = R + n_IR;
    yn[3] = cases_cumul + n_SI;
    yn[4] = cases_inc + n_SI;
  }
  static double compare_data(double, const double* y, const data_type& d,
                             const shared_state& sh, rng_state& rng) {
    if (std::isnan(d.incidence)) {
      return 0;
    }
    const double noise = exponential_rate(rng, sh.exp_noise);
    const double lambda = y[4] + noise;
    return dpois_log(d.incidence, lambda);
  }

  // Discrete adjoint (sir.cpp:126-198).  n_adjoint = 5 states + 3
  // parameters (beta, gamma, I0).  `exact` false: the reference's
  // hand-written adjoint term for term -- its gamma term multiplies by
  // exp(-gamma) (sir.cpp:175; the derivative of p_IR = 1 - exp(-gamma dt)
  // is dt exp(-gamma dt), equal only at dt = 1), its adjoint_initial adds
  // adj_I but not the -adj_S of S0 = N - I0 to the I0 gradient
  // (sir.cpp:126-141), and its compare adjoint has no missing-data branch
  // (sir.cpp:179-197: NaN data give a NaN gradient).  `exact` true: the
  // derivative of the likelihood the forward run computes, which the
  // finite-difference checks in the tests settle.
  static constexpr size_t n_gradient = 3;
  static void adjoint_initial(const double*, const double* adj, const shared_state&,
                              double* adj_next, bool exact) {
    for (size_t i = 0; i < 8; ++i) adj_next[i] = adj[i];
    adj_next[7] = exact ? adj[7] + adj[1] - adj[0] : adj[7] + adj[1];
  }
  static void adjoint_update(double dt, const double* y, const double* adj,
                             const shared_state& sh, double* an, bool exact) {
    const double S = y[0];
    const double I = y[1];
    const double e_SI = std::exp(-sh.beta * I * dt / sh.N);
    // the reference evaluates p_SI in the forward model's association
    // (sir.cpp:161)
    const double p_SI = exact ? 1 - e_SI : 1 - std::exp(-sh.beta * I / sh.N * dt);
    const double e_IR = std::exp(-sh.gamma * dt);
    const double p_IR = 1 - e_IR;
    const double adj_n_IR = -adj[1] + adj[2];
    const double adj_n_SI = adj[4] + adj[3] + adj[1] - adj[0];
    const double adj_p_IR = adj_n_IR * I;
    const double adj_p_SI = adj_n_SI * S;
    an[0] = adj_n_SI * p_SI + adj[0];
    an[1] = adj_n_IR * p_IR + adj_p_SI * (sh.beta * dt) * e_SI / sh.N + adj[1];
    an[2] = adj[2];
    an[3] = adj[3];
    an[4] = adj[4];
    an[5] = adj[5] + adj_p_SI * I * dt * e_SI / sh.N;
    an[6] = exact ? adj[6] + adj_p_IR * dt * e_IR : adj[6] + adj_p_IR * std::exp(-sh.gamma);
    an[7] = adj[7];
  }
  static void adjoint_compare_data(const double* y, const double* adj,
                                   const data_type& d, const shared_state& sh,
                                   double* an, bool exact) {
    for (size_t i = 0; i < 8; ++i) an[i] = adj[i];
    if (exact && std::isnan(d.incidence)) {
      return;
    }
    const double lambda = y[4] + 1 / sh.exp_noise;
    an[4] = adj[4] + (d.incidence / lambda - 1);
  }
};

// inst/examples/sirode.cpp:9-102
struct sirode {
  static constexpr size_t n_state = 5;
  static constexpr size_t n_pars = 4;
  struct shared_state {
    double N, I0, beta, gamma;
  };
  struct data_type {
    double incidence;
  };
  static shared_state build_shared(const double* p) {
    return shared_state{p[0], p[1], p[2], p[3]};
  }
  static constexpr size_t n_data = 1;
  static data_type build_data(const double* d) { return data_type{d[0]}; }
  static constexpr double zero_period = 1;
  static constexpr size_t zero_first = 4;
  static constexpr size_t zero_count = 1;
  static void initial(double, const shared_state& sh, rng_state&, double* y) {
    y[0] = sh.N - sh.I0;
    y[1] = sh.I0;
    y[2] = 0;
    y[3] = 0;
    y[4] = 0;
  }
  static void rhs(double, const double* y, const shared_state& sh, double* dy) {
    const double S = y[0];
    const double I = y[1];
    const double rate_SI = sh.beta * S * I / sh.N;
    const double rate_IR = sh.gamma * I;
    dy[0] = -rate_SI;
    dy[1] = rate_SI - rate_IR;
    dy[2] = rate_IR;
    dy[3] = rate_SI;
    dy[4] = rate_SI;
  }
  static double compare_data(double, const double* y, const data_type& d,
                             const shared_state&, rng_state&) {
    if (std::isnan(d.incidence)) {
      return 0;
    }
    return dpois_log(d.incidence, y[4]);
  }
};

// Forward sensitivities of the SIR ODE with respect to one parameter
// (beta, gamma or I0): the state is augmented with s = dy/dtheta, which
// obeys ds/dt = J(y) s + df/dtheta.  No counterpart in the reference (its
// only gradient is the discrete adjoint).  Layout [S, I, R, cum, sS, sI,
// sR, scum, inc, sinc] keeps the two reset variables contiguous; only the
// five model states enter the error norm and the initial step, so the
// step sequence, the states and the likelihood are those of `sirode`.
struct sirode_sens {
  static constexpr size_t n_state = 10;
  struct shared_state {
    sirode::shared_state base;
    int param;  // 0: beta, 1: gamma, 2: I0
  };
  static constexpr double zero_period = 1;
  static constexpr size_t zero_first = 8;
  static constexpr size_t zero_count = 2;
  static bool error_component(size_t i) { return i < 4 || i == 8; }
  static constexpr size_t n_error = 5;
  static void initial(const shared_state& sh, double* y) {
    for (size_t i = 0; i < n_state; ++i) y[i] = 0;
    y[0] = sh.base.N - sh.base.I0;
    y[1] = sh.base.I0;
    if (sh.param == 2) {
      y[4] = -1;
      y[5] = 1;
    }
  }
  static void rhs(double, const double* y, const shared_state& sh, double* dy) {
    const auto& b = sh.base;
    const double S = y[0];
    const double I = y[1];
    const double rate_SI = b.beta * S * I / b.N;
    const double rate_IR = b.gamma * I;
    dy[0] = -rate_SI;
    dy[1] = rate_SI - rate_IR;
    dy[2] = rate_IR;
    dy[3] = rate_SI;
    dy[8] = rate_SI;
    const double sS = y[4];
    const double sI = y[5];
    double d_SI = b.beta * (sS * I + S * sI) / b.N;
    double d_IR = b.gamma * sI;
    if (sh.param == 0) d_SI += S * I / b.N;
    if (sh.param == 1) d_IR += I;
    dy[4] = -d_SI;
    dy[5] = d_SI - d_IR;
    dy[6] = d_IR;
    dy[7] = d_SI;
    dy[9] = d_SI;
  }
};

// Random walk, inst/examples/walk.cpp:1-100, with the walker count `len`
// fixed at compile time (the reference reads it from the parameters as a
// constant).  Parameter row: sd, len, random_initial.  No data.
template <size_t LEN>
struct walk {
  static constexpr size_t n_state = LEN;
  static constexpr size_t n_pars = 3;
  struct shared_state {
    double sd;
    bool random_initial;
  };
  struct data_type {};
  static shared_state build_shared(const double* p) { return shared_state{p[0], p[2] != 0}; }
  static constexpr size_t n_data = 0;
  static data_type build_data(const double*) { return data_type{}; }
  static constexpr double zero_period = 1;
  static constexpr size_t zero_first = 0;
  static constexpr size_t zero_count = 0;
  // walk.cpp:50-62
  static void initial(double, const shared_state& sh, rng_state& rng, double* y) {
    for (size_t i = 0; i < LEN; ++i) {
      y[i] = sh.random_initial ? normal(rng, 0, sh.sd) : 0;
    }
  }
  // walk.cpp:65-76
  static void update(double, double dt, const double* y, const shared_state& sh,
                     rng_state& rng, double* yn) {
    for (size_t i = 0; i < LEN; ++i) yn[i] = normal(rng, y[i], sh.sd * dt);
  }
  static double compare_data(double, const double*, const data_type&, const shared_state&,
                             rng_state&) {
    return 0;
  }
};

// Age-structured stochastic SEIR with 16 age groups.  No counterpart exists
// in the reference; written to the same contract.  State layout:
// S[0:16] E[16:32] I[32:48] R[48:64] inc[64:80]; incidence is reset every
// 7 days (weekly observations of each age group's incidence).
// Force of infection lambda_a = beta * sum_b C[a][b] I_b / N_b.
struct seir16 {
  static constexpr size_t n_age = 16;
  static constexpr size_t n_state = 5 * n_age;
  // beta, sigma, gamma, exp_noise, I0 (seeded into every age group),
  // N[16], C[16*16] row-major
  static constexpr size_t n_pars = 5 + n_age + n_age * n_age;
  struct shared_state {
    double beta, sigma, gamma, exp_noise, I0;
    double N[n_age];
    double C[n_age * n_age];
  };
  struct data_type {
    double incidence[n_age];
  };
  static shared_state build_shared(const double* p) {
    shared_state sh;
    sh.beta = p[0];
    sh.sigma = p[1];
    sh.gamma = p[2];
    sh.exp_noise = p[3];
    sh.I0 = p[4];
    for (size_t a = 0; a < n_age; ++a) sh.N[a] = p[5 + a];
    for (size_t i = 0; i < n_age * n_age; ++i) sh.C[i] = p[5 + n_age + i];
    return sh;
  }
  static constexpr size_t n_data = n_age;
  static data_type build_data(const double* d) {
    data_type ret;
    for (size_t a = 0; a < n_age; ++a) ret.incidence[a] = d[a];
    return ret;
  }
  static constexpr double zero_period = 7;
  static constexpr size_t zero_first = 4 * n_age;
  static constexpr size_t zero_count = n_age;

  static void initial(double, const shared_state& sh, rng_state&, double* y) {
    for (size_t a = 0; a < n_age; ++a) {
      y[a] = sh.N[a] - sh.I0;
      y[n_age + a] = 0;
      y[2 * n_age + a] = sh.I0;
      y[3 * n_age + a] = 0;
      y[4 * n_age + a] = 0;
    }
  }
  static void update(double, double dt, const double* y, const shared_state& sh,
                     rng_state& rng, double* yn) {
    const double* S = y;
    const double* E = y + n_age;
    const double* I = y + 2 * n_age;
    const double* R = y + 3 * n_age;
    const double* inc = y + 4 * n_age;
    double x[n_age];
    for (size_t b = 0; b < n_age; ++b) x[b] = I[b] / sh.N[b];
    const double p_EI = 1 - std::exp(-sh.sigma * dt);
    const double p_IR = 1 - std::exp(-sh.gamma * dt);
    for (size_t a = 0; a < n_age; ++a) {
      double foi = 0;
      for (size_t b = 0; b < n_age; ++b) foi += sh.C[a * n_age + b] * x[b];
      const double p_SE = 1 - std::exp(-sh.beta * foi * dt);
      const double n_SE = binomial(rng, S[a], p_SE);
      const double n_EI = binomial(rng, E[a], p_EI);
      const double n_IR = binomial(rng, I[a], p_IR);
      yn[a] = S[a] - n_SE;
      yn[n_age + a] = E[a] + n_SE - n_EI;
      yn[2 * n_age + a] = I[a] + n_EI - n_IR;
      yn[3 * n_age + a] = R[a] + n_IR;
      yn[4 * n_age + a] = inc[a] + n_SE;
    }
  }
  static double compare_data(double, const double* y, const data_type& d,
                             const shared_state& sh, rng_state& rng) {
    double ll = 0;
    for (size_t a = 0; a < n_age; ++a) {
      if (std::isnan(d.incidence[a])) {
        continue;
      }
      const double noise = exponential_rate(rng, sh.exp_noise);
      ll += dpois_log(d.incidence[a], y[4 * n_age + a] + noise);
    }
    return ll;
  }
};

// ---------------------------------------------------------------------------
// dust2/filter_details.hpp:11-74
// ---------------------------------------------------------------------------
inline double scale_log_weights(size_t n, double* w) {
  if (n == 1) {
    return w[0];
  }
  const double neg_inf = -std::numeric_limits<double>::infinity();
  double max_w = neg_inf;
  for (size_t i = 0; i < n; ++i) {
    if (std::isnan(w[i])) {
      w[i] = neg_inf;
    } else {
      max_w = std::max(max_w, w[i]);
    }
  }
  if (max_w == neg_inf) {
    return max_w;
  }
  double tot = 0;
  for (size_t i = 0; i < n; ++i) {
    w[i] = std::exp(w[i] - max_w);
    tot += w[i];
  }
  return std::log(tot / n) + max_w;
}

// Systematic resampling: the ancestor of slot i is the first j whose
// running (sequential, left to right) sum reaches tot*(u + i)/n, minus one.
template <typename Index>
void resample_weight(size_t n, const double* w, double u, Index* idx) {
  double tot = 0;
  for (size_t i = 0; i < n; ++i) tot += w[i];
  const double uu0 = tot * u / n;
  const double du = tot / n;
  double ww = 0;
  size_t j = 0;
  for (size_t i = 0; i < n; ++i) {
    const double uu = uu0 + i * du;
    while (ww < uu && j < n) {
      ww += w[j];
      ++j;
    }
    idx[i] = static_cast<Index>(j == 0 ? 0 : j - 1);
  }
}

// ---------------------------------------------------------------------------
// Discrete-time particle container (dust2/discrete/system.hpp:20-483)
// ---------------------------------------------------------------------------
template <typename T>
class discrete_system {
 public:
  using shared_state = typename T::shared_state;
  using data_type = typename T::data_type;

  discrete_system(std::vector<shared_state> shared, double time, double dt,
                  size_t n_particles, const std::vector<uint64_t>& seed,
                  bool deterministic, int n_threads)
      : n_particles_(n_particles),
        n_groups_(shared.size()),
        n_total_(n_particles * shared.size()),
        state_(T::n_state * n_total_, 0.0),
        state_next_(T::n_state * n_total_, 0.0),
        shared_(std::move(shared)),
        time_(time),
        dt_(dt),
        errors_(n_total_),
        rng_(prng(n_total_, seed, deterministic)),
        n_threads_(n_threads) {}

  void run_to_time(double time, const std::vector<size_t>& index_group) {
    const size_t n_steps = std::round(std::max(0.0, time - time_) / dt_);
    double* state = state_.data();
    double* state_next = state_next_.data();
    const size_t n_index = index_group.size();
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(n_threads_) collapse(2)
#endif
    for (size_t ii = 0; ii < n_index; ++ii) {
      for (size_t j = 0; j < n_particles_; ++j) {
        const size_t i = index_group[ii];
        const size_t k = n_particles_ * i + j;
        try {
          run_particle(time_, dt_, n_steps, shared_[i],
                       state + k * T::n_state, rng_[k],
                       state_next + k * T::n_state);
        } catch (const std::exception& e) {
          errors_.capture(e, k);
        }
      }
    }
    errors_.report();
    if (n_steps % 2 == 1) {
      std::swap(state_, state_next_);
    }
    time_ = time_ + n_steps * dt_;
  }

  void set_state_initial(const std::vector<size_t>& index_group) {
    errors_.reset();
    const size_t n_index = index_group.size();
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(n_threads_) collapse(2)
#endif
    for (size_t ii = 0; ii < n_index; ++ii) {
      for (size_t j = 0; j < n_particles_; ++j) {
        const size_t i = index_group[ii];
        const size_t k = n_particles_ * i + j;
        T::initial(time_, shared_[i], rng_[k], state_.data() + k * T::n_state);
      }
    }
  }

  void set_state_initial() {
    std::vector<size_t> all(n_groups_);
    std::iota(all.begin(), all.end(), 0);
    set_state_initial(all);
  }

  void set_state(const double* src) {
    errors_.reset();
    std::copy_n(src, state_.size(), state_.begin());
  }

  void compare_data(const data_type* data, size_t n_groups_data,
                    const std::vector<size_t>& index_group, double* out) {
    const bool shared_data = n_groups_data == 1;
    const size_t n_index = index_group.size();
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(n_threads_) collapse(2)
#endif
    for (size_t ii = 0; ii < n_index; ++ii) {
      for (size_t j = 0; j < n_particles_; ++j) {
        const size_t i = index_group[ii];
        const size_t k = n_particles_ * i + j;
        const data_type& d = data[shared_data ? 0 : i];
        try {
          out[k] = T::compare_data(time_, state_.data() + k * T::n_state, d,
                                   shared_[i], rng_[k]);
        } catch (const std::exception& e) {
          errors_.capture(e, k);
        }
      }
    }
    errors_.report();
  }

  // Serial gather (dust2/discrete/system.hpp:169-181).  RNG streams stay
  // with their slot.
  template <typename Index>
  void reorder(const Index* idx, const std::vector<size_t>& index_group) {
    for (size_t i : index_group) {
      for (size_t j = 0; j < n_particles_; ++j) {
        const size_t k_to = n_particles_ * i + j;
        const size_t k_from = n_particles_ * i + idx[k_to];
        std::copy_n(state_.begin() + k_from * T::n_state, T::n_state,
                    state_next_.begin() + k_to * T::n_state);
      }
    }
    std::swap(state_, state_next_);
  }

  // History-saving run for the adjoint (dust2/discrete/system.hpp:94-119,
  // 425-451): history slot s holds the state after s steps, before any
  // zeroing applied at the start of step s.
  void run_to_time_history(double time, double* history) {
    const size_t n_steps = std::round(std::max(0.0, time - time_) / dt_);
    const size_t stride = T::n_state * n_total_;
    std::copy_n(state_.begin(), stride, history);
    for (size_t k = 0; k < n_total_; ++k) {
      const size_t i = k / n_particles_;
      double* model = state_.data() + k * T::n_state;
      double* curr = model;
      double* next = history + stride + k * T::n_state;
      for (size_t s = 0; s < n_steps; ++s, next += stride) {
        const double time_s = time_ + s * dt_;
        if (std::fmod(time_s, T::zero_period) == 0) {
          std::copy_n(curr, T::n_state, model);
          curr = model;
          for (size_t z = 0; z < T::zero_count; ++z) curr[T::zero_first + z] = 0;
        }
        T::update(time_s, dt_, curr, shared_[i], rng_[k], next);
        curr = next;
      }
      std::copy_n(curr, T::n_state, model);
    }
    time_ = time_ + n_steps * dt_;
  }

  static void run_particle(double time, double dt, size_t n_steps,
                           const shared_state& shared, double* state,
                           rng_state& rng, double* state_next) {
    for (size_t s = 0; s < n_steps; ++s) {
      const double time_s = time + s * dt;
      if (std::fmod(time_s, T::zero_period) == 0) {
        for (size_t z = 0; z < T::zero_count; ++z) state[T::zero_first + z] = 0;
      }
      T::update(time_s, dt, state, shared, rng, state_next);
      std::swap(state, state_next);
    }
  }

  void set_time(double t) { time_ = t; }
  double time() const { return time_; }
  double dt() const { return dt_; }
  size_t n_particles() const { return n_particles_; }
  size_t n_groups() const { return n_groups_; }
  std::vector<double>& state() { return state_; }
  std::vector<rng_state>& rng() { return rng_; }
  std::vector<shared_state>& shared() { return shared_; }

 private:
  size_t n_particles_;
  size_t n_groups_;
  size_t n_total_;
  std::vector<double> state_;
  std::vector<double> state_next_;
  std::vector<shared_state> shared_;
  double time_;
  double dt_;
  errors errors_;
  std::vector<rng_state> rng_;
  int n_threads_;
};

// ---------------------------------------------------------------------------
// Bootstrap particle filter (dust2/filter.hpp:86-254)
// ---------------------------------------------------------------------------
template <typename T>
class filter {
 public:
  using data_type = typename T::data_type;
  discrete_system<T> sys;

  filter(discrete_system<T> sys_, double time_start, std::vector<double> time,
         std::vector<data_type> data, const std::vector<uint64_t>& seed_filter)
      : sys(std::move(sys_)),
        time_start_(time_start),
        time_(std::move(time)),
        data_(std::move(data)),
        n_groups_data_(data_.size() / time_.size()),
        rng_(prng(2 * sys.n_groups(), seed_filter, false)),
        ll_(sys.n_groups(), 0),
        ll_step_(sys.n_groups() * sys.n_particles(), 0),
        index_(sys.n_groups() * sys.n_particles()) {}

  // `ancestry`, when given, receives the n_time x G x P ancestor indices.
  // ancestry [T][G][P] and traj [T][G][P][S] (the states after each
  // time's reorder, trajectories_.add of dust2/filter.hpp:89-91), when
  // given, for the groups of index_group
  void run(bool set_initial, const std::vector<size_t>& index_group,
           uint32_t* ancestry = nullptr, double* traj = nullptr) {
    const size_t P = sys.n_particles();
    std::fill(ll_.begin(), ll_.end(), 0);
    sys.set_time(time_start_);
    if (set_initial) {
      std::vector<size_t> all(sys.n_groups());
      std::iota(all.begin(), all.end(), 0);
      sys.set_state_initial(all);
    }
    const data_type* it_data = data_.data();
    for (size_t t = 0; t < time_.size(); ++t) {
      sys.run_to_time(time_[t], index_group);
      sys.compare_data(it_data, n_groups_data_, index_group, ll_step_.data());
      it_data += n_groups_data_;
      for (size_t i : index_group) {
        double* w = ll_step_.data() + i * P;
        size_t* idx = index_.data() + i * P;
        const bool any = std::any_of(w, w + P, [](double v) { return v != 0; });
        if (any) {
          ll_[i] += scale_log_weights(P, w);
          const double u = random_real(rng_[2 * i]);
          resample_weight(P, w, u, idx);
        } else {
          for (size_t j = 0; j < P; ++j) idx[j] = j;
        }
      }
      sys.reorder(index_.data(), index_group);
      if (ancestry) {
        for (size_t i : index_group) {
          for (size_t j = 0; j < P; ++j) {
            ancestry[(t * sys.n_groups() + i) * P + j] =
                static_cast<uint32_t>(index_[i * P + j]);
          }
        }
      }
      if (traj) {
        const size_t block = P * T::n_state;
        for (size_t i : index_group) {
          std::copy(sys.state().begin() + i * block, sys.state().begin() + (i + 1) * block,
                    traj + (t * sys.n_groups() + i) * block);
        }
      }
    }
  }

  const std::vector<double>& log_likelihood() const { return ll_; }
  std::vector<rng_state>& rng() { return rng_; }

 private:
  double time_start_;
  std::vector<double> time_;
  std::vector<data_type> data_;
  size_t n_groups_data_;
  std::vector<rng_state> rng_;
  std::vector<double> ll_;
  std::vector<double> ll_step_;
  std::vector<size_t> index_;
};

// ---------------------------------------------------------------------------
// Dormand-Prince 5(4) with Shampine dense output and a PI controller
// (dust2/continuous/solver.hpp:75-545, control.hpp:11-39)
// ---------------------------------------------------------------------------
struct ode_control {
  size_t max_steps = 10000;
  double atol = 1e-6;
  double rtol = 1e-6;
  double step_size_min = std::numeric_limits<double>::epsilon();
  double step_size_max = std::numeric_limits<double>::infinity();
  double factor_safe = 0.9;
  double factor_min = 0.2;
  double factor_max = 10.0;
  double beta = 0.04;
  double constant = 0.2 - 0.04 * 0.75;
};

template <size_t N>
struct ode_internals {
  // last accepted step: interval and dense-output coefficients
  double t0 = 0, t1 = 0, h = 0;
  double c1[N] = {}, c2[N] = {}, c3[N] = {}, c4[N] = {}, c5[N] = {};
  double dydt[N] = {};
  double step_size = 0;
  double error = 0;
  size_t n_steps = 0;
  size_t n_steps_accepted = 0;
  size_t n_steps_rejected = 0;
  void reset() {
    step_size = 0;
    error = 0;
    n_steps = 0;
    n_steps_accepted = 0;
    n_steps_rejected = 0;
  }
  double interpolate(double t, size_t i) const {
    const double u = (t - t0) / h;
    const double v = 1 - u;
    return c1[i] + u * (c2[i] + v * (c3[i] + u * (c4[i] + v * c5[i])));
  }
};

namespace dp5 {
constexpr double C2 = 0.2, C3 = 0.3, C4 = 0.8, C5 = 8.0 / 9.0;
constexpr double A21 = 0.2;
constexpr double A31 = 3.0 / 40.0, A32 = 9.0 / 40.0;
constexpr double A41 = 44.0 / 45.0, A42 = -56.0 / 15.0, A43 = 32.0 / 9.0;
constexpr double A51 = 19372.0 / 6561.0, A52 = -25360.0 / 2187.0,
                 A53 = 64448.0 / 6561.0, A54 = -212.0 / 729.0;
constexpr double A61 = 9017.0 / 3168.0, A62 = -355.0 / 33.0,
                 A63 = 46732.0 / 5247.0, A64 = 49.0 / 176.0,
                 A65 = -5103.0 / 18656.0;
constexpr double A71 = 35.0 / 384.0, A73 = 500.0 / 1113.0, A74 = 125.0 / 192.0,
                 A75 = -2187.0 / 6784.0, A76 = 11.0 / 84.0;
constexpr double E1 = 71.0 / 57600.0, E3 = -71.0 / 16695.0, E4 = 71.0 / 1920.0,
                 E5 = -17253.0 / 339200.0, E6 = 22.0 / 525.0, E7 = -1.0 / 40.0;
constexpr double D1 = -12715105075.0 / 11282082432.0,
                 D3 = 87487479700.0 / 32700410799.0,
                 D4 = -10690763975.0 / 1880347072.0,
                 D5 = 701980252875.0 / 199316789632.0,
                 D6 = -1453857185.0 / 822651844.0,
                 D7 = 69997945.0 / 29380423.0;
}  // namespace dp5

// Models may restrict the error norm and the initial step to some state
// components (error_component(i), n_error); by default all take part.
template <typename U, typename = void>
struct has_error_mask : std::false_type {};
template <typename U>
struct has_error_mask<U, std::void_t<decltype(U::error_component(0))>> : std::true_type {};

template <typename T>
class ode_solver {
 public:
  static constexpr size_t N = T::n_state;
  static bool in_error(size_t i) {
    if constexpr (has_error_mask<T>::value) {
      return T::error_component(i);
    } else {
      (void)i;
      return true;
    }
  }
  static constexpr size_t n_error() {
    if constexpr (has_error_mask<T>::value) {
      return T::n_error;
    } else {
      return N;
    }
  }
  using shared_state = typename T::shared_state;
  using internals = ode_internals<N>;
  ode_control control;

  explicit ode_solver(ode_control c) : control(c) {}

  double try_step(double t, double h, const double* y, const double* k1,
                  double* c5, const shared_state& sh) {
    using namespace dp5;
    const double t_next = t + h;
    for (size_t i = 0; i < N; ++i) y_next_[i] = y[i] + h * A21 * k1[i];
    T::rhs(t + C2 * h, y_next_, sh, k2_);
    for (size_t i = 0; i < N; ++i)
      y_next_[i] = y[i] + h * (A31 * k1[i] + A32 * k2_[i]);
    T::rhs(t + C3 * h, y_next_, sh, k3_);
    for (size_t i = 0; i < N; ++i)
      y_next_[i] = y[i] + h * (A41 * k1[i] + A42 * k2_[i] + A43 * k3_[i]);
    T::rhs(t + C4 * h, y_next_, sh, k4_);
    for (size_t i = 0; i < N; ++i)
      y_next_[i] = y[i] + h * (A51 * k1[i] + A52 * k2_[i] + A53 * k3_[i] +
                               A54 * k4_[i]);
    T::rhs(t + C5 * h, y_next_, sh, k5_);
    for (size_t i = 0; i < N; ++i)
      y_stiff_[i] = y[i] + h * (A61 * k1[i] + A62 * k2_[i] + A63 * k3_[i] +
                                A64 * k4_[i] + A65 * k5_[i]);
    T::rhs(t_next, y_stiff_, sh, k6_);
    for (size_t i = 0; i < N; ++i)
      y_next_[i] = y[i] + h * (A71 * k1[i] + A73 * k3_[i] + A74 * k4_[i] +
                               A75 * k5_[i] + A76 * k6_[i]);
    T::rhs(t_next, y_next_, sh, k2_);
    for (size_t i = 0; i < N; ++i)
      c5[i] = h * (D1 * k1[i] + D3 * k3_[i] + D4 * k4_[i] + D5 * k5_[i] +
                   D6 * k6_[i] + D7 * k2_[i]);
    for (size_t i = 0; i < N; ++i)
      k4_[i] = h * (E1 * k1[i] + E3 * k3_[i] + E4 * k4_[i] + E5 * k5_[i] +
                    E6 * k6_[i] + E7 * k2_[i]);
    double err = 0;
    for (size_t i = 0; i < N; ++i) {
      if (!in_error(i)) continue;
      const double sk =
          k4_[i] / (control.atol +
                    control.rtol * std::max(std::fabs(y[i]), std::fabs(y_next_[i])));
      err += sk * sk;
    }
    return std::sqrt(err / n_error());
  }

  double step(double t, double t_end, double* y, internals& in,
              const shared_state& sh) {
    const double facc1 = 1 / control.factor_min;
    const double facc2 = 1 / control.factor_max;
    bool success = false;
    bool reject = false;
    bool truncated = false;
    double h = in.step_size;
    while (!success) {
      if (in.n_steps > control.max_steps) {
        throw std::runtime_error("too many steps");
      }
      if (h < control.step_size_min) {
        throw std::runtime_error("step too small");
      }
      if (h <= std::fabs(t) * std::numeric_limits<double>::epsilon()) {
        throw std::runtime_error("step size vanished");
      }
      truncated = t + h > t_end;
      if (truncated) {
        h = t_end - t;
      }
      const double err = try_step(t, h, y, in.dydt, in.c5, sh);
      in.n_steps++;
      const double fac11 = std::pow(err, control.constant);
      if (err <= 1) {
        success = true;
        // dense output (solver.hpp:335-352)
        in.t0 = t;
        in.t1 = t + h;
        in.h = h;
        for (size_t i = 0; i < N; ++i) {
          const double ydiff = y_next_[i] - y[i];
          const double bspl = h * in.dydt[i] - ydiff;
          in.c1[i] = y[i];
          in.c2[i] = ydiff;
          in.c3[i] = bspl;
          in.c4[i] = -h * k2_[i] + ydiff - bspl;
        }
        // accept (solver.hpp:354-357)
        for (size_t i = 0; i < N; ++i) {
          in.dydt[i] = k2_[i];
          y[i] = y_next_[i];
        }
        in.n_steps_accepted++;
        if (!truncated) {
          const double fac_old = std::max(in.error, 1e-4);
          double fac = fac11 / std::pow(fac_old, control.beta);
          fac = std::max(std::min(fac / control.factor_safe, facc1), facc2);
          const double h_new = h / fac;
          const double h_max = reject ? h : control.step_size_max;
          in.step_size = std::min(h_new, h_max);
          in.error = err;
        }
      } else {
        reject = true;
        if (in.n_steps_accepted >= 1) {
          in.n_steps_rejected++;
        }
        h /= std::min(facc1, fac11 / control.factor_safe);
      }
    }
    return truncated ? t_end : t + h;
  }

  // Zeroing of reset variables at the start of a step (solver.hpp:437-465).
  // When the boundary falls inside the last step, the value interpolated at
  // the boundary is subtracted.
  void apply_zero_every(double t, double* y, const internals& in) {
    if (in.h == 0) {
      return;
    }
    const double period = T::zero_period;
    const int n = std::floor(t / period);
    const int m = std::floor(in.t0 / period);
    if (n > m) {
      const double t_reset = n * period;
      for (size_t z = 0; z < T::zero_count; ++z) {
        const size_t j = T::zero_first + z;
        if (t_reset == t) {
          y[j] = 0;
        } else {
          y[j] -= in.interpolate(t_reset, j);
        }
      }
    }
  }

  void apply_zero_every_final(double t, double* y, const internals& in) {
    if (in.h == 0) {
      return;
    }
    const double period = T::zero_period;
    if (in.h > period) {
      const int n = std::ceil(t / period);
      const int m = std::ceil(in.t0 / period);
      if (n > m) {
        const double t_reset = (n - 1) * period;
        for (size_t z = 0; z < T::zero_count; ++z) {
          const size_t j = T::zero_first + z;
          y[j] -= in.interpolate(t_reset, j);
        }
      }
    }
  }

  void run(double t, double t_end, double* y, internals& in,
           const shared_state& sh) {
    while (t < t_end) {
      apply_zero_every(t, y, in);
      t = step(t, t_end, y, in, sh);
    }
    apply_zero_every_final(t, y, in);
  }

  // Hairer's initial step size (solver.hpp:264-324)
  void initialise(double t, const double* y, internals& in,
                  const shared_state& sh) {
    in.reset();
    double* f0 = in.dydt;
    double* f1 = k3_;
    double* y1 = y_next_;
    T::rhs(t, y, sh, f0);
    double norm_f = 0;
    double norm_y = 0;
    for (size_t i = 0; i < N; ++i) {
      if (!in_error(i)) continue;
      const double sk = control.atol + control.rtol * std::fabs(y[i]);
      norm_f += (f0[i] / sk) * (f0[i] / sk);
      norm_y += (y[i] / sk) * (y[i] / sk);
    }
    double h = (norm_f <= 1e-10 || norm_y <= 1e-10)
                   ? 1e-6
                   : std::sqrt(norm_y / norm_f) * 0.01;
    h = std::min(h, control.step_size_max);
    for (size_t i = 0; i < N; ++i) y1[i] = y[i] + h * f0[i];
    T::rhs(t + h, y1, sh, f1);
    double der2 = 0;
    for (size_t i = 0; i < N; ++i) {
      if (!in_error(i)) continue;
      const double sk = control.atol + control.rtol * std::fabs(y[i]);
      der2 += ((f1[i] - f0[i]) / sk) * ((f1[i] - f0[i]) / sk);
    }
    der2 = std::sqrt(der2) / h;
    const double der12 = std::max(std::fabs(der2), std::sqrt(norm_f));
    const double h1 = (der12 <= 1e-15) ? std::max(1e-6, std::fabs(h) * 1e-3)
                                       : std::pow(0.01 / der12, 1.0 / 5);
    h = std::min(100 * std::fabs(h), h1);
    if (!std::isfinite(h)) {
      throw std::runtime_error("Initial step size was not finite");
    }
    h = std::max(h, control.step_size_min * 100);
    h = std::min(h, control.step_size_max);
    in.step_size = h;
  }

 private:
  double y_next_[N];
  double y_stiff_[N];
  double k2_[N], k3_[N], k4_[N], k5_[N], k6_[N];
};

// Continuous-time container (dust2/continuous/system.hpp:24-586), pure-ODE
// path without events, delays or output variables.
template <typename T>
class continuous_system {
 public:
  using shared_state = typename T::shared_state;
  using data_type = typename T::data_type;
  using internals = ode_internals<T::n_state>;

  continuous_system(std::vector<shared_state> shared, double time,
                    size_t n_particles, ode_control control, int n_threads)
      : n_particles_(n_particles),
        n_groups_(shared.size()),
        n_total_(n_particles * shared.size()),
        state_(T::n_state * n_total_, 0.0),
        internals_(n_total_),
        shared_(std::move(shared)),
        time_(time),
        control_(control),
        errors_(n_total_),
        rng_(n_total_, rng_state{{0, 0, 0, 0}, true}),
        requires_initialise_(n_groups_, 1),
        n_threads_(n_threads) {}

  void initialise_solver(const std::vector<size_t>& index_group) {
    bool any = false;
    for (size_t i : index_group) any = any || requires_initialise_[i];
    if (!any) {
      return;
    }
    errors_.reset();
    const size_t n_index = index_group.size();
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(n_threads_) collapse(2)
#endif
    for (size_t ii = 0; ii < n_index; ++ii) {
      for (size_t j = 0; j < n_particles_; ++j) {
        const size_t i = index_group[ii];
        if (!requires_initialise_[i]) continue;
        const size_t k = n_particles_ * i + j;
        try {
          ode_solver<T> solver(control_);
          solver.initialise(time_, state_.data() + k * T::n_state, internals_[k],
                            shared_[i]);
        } catch (const std::exception& e) {
          errors_.capture(e, k);
        }
      }
    }
    errors_.report();
    for (size_t i : index_group) requires_initialise_[i] = 0;
  }

  void run_to_time(double time, const std::vector<size_t>& index_group) {
    initialise_solver(index_group);
    const size_t n_index = index_group.size();
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(n_threads_) collapse(2)
#endif
    for (size_t ii = 0; ii < n_index; ++ii) {
      for (size_t j = 0; j < n_particles_; ++j) {
        const size_t i = index_group[ii];
        const size_t k = n_particles_ * i + j;
        try {
          ode_solver<T> solver(control_);
          solver.run(time_, time, state_.data() + k * T::n_state, internals_[k],
                     shared_[i]);
        } catch (const std::exception& e) {
          errors_.capture(e, k);
        }
      }
    }
    errors_.report();
    time_ = time;
  }

  void set_state_initial() {
    errors_.reset();
    for (size_t k = 0; k < n_total_; ++k) {
      T::initial(time_, shared_[k / n_particles_], rng_[k],
                 state_.data() + k * T::n_state);
    }
    std::fill(requires_initialise_.begin(), requires_initialise_.end(), 1);
  }

  void set_time(double t) {
    time_ = t;
    std::fill(requires_initialise_.begin(), requires_initialise_.end(), 1);
  }

  void compare_data(const data_type* data, size_t n_groups_data,
                    const std::vector<size_t>& index_group, double* out) {
    const bool shared_data = n_groups_data == 1;
    for (size_t i : index_group) {
      for (size_t j = 0; j < n_particles_; ++j) {
        const size_t k = n_particles_ * i + j;
        out[k] = T::compare_data(time_, state_.data() + k * T::n_state,
                                 data[shared_data ? 0 : i], shared_[i], rng_[k]);
      }
    }
  }

  std::vector<double>& state() { return state_; }
  std::vector<internals>& ode_internals_() { return internals_; }
  size_t n_particles() const { return n_particles_; }
  size_t n_groups() const { return n_groups_; }

 private:
  size_t n_particles_;
  size_t n_groups_;
  size_t n_total_;
  std::vector<double> state_;
  std::vector<internals> internals_;
  std::vector<shared_state> shared_;
  double time_;
  ode_control control_;
  errors errors_;
  std::vector<rng_state> rng_;
  std::vector<char> requires_initialise_;
  int n_threads_;
};

// Deterministic likelihood (dust2/unfilter.hpp:42-72).  Data advance by
// n_groups_data per time; with per-group data that is the reference's
// n_groups stride (unfilter.hpp:52).
template <typename Sys>
void unfilter_run(Sys& sys, double time_start, const std::vector<double>& time,
                  const typename Sys::data_type* data, size_t n_groups_data,
                  double* ll) {
  const size_t n_total = sys.n_groups() * sys.n_particles();
  std::vector<size_t> all(sys.n_groups());
  std::iota(all.begin(), all.end(), 0);
  std::vector<double> ll_step(n_total);
  std::fill(ll, ll + n_total, 0.0);
  sys.set_time(time_start);
  sys.set_state_initial();
  for (size_t t = 0; t < time.size(); ++t) {
    sys.run_to_time(time[t], all);
    sys.compare_data(data + t * n_groups_data, n_groups_data, all, ll_step.data());
    for (size_t k = 0; k < n_total; ++k) ll[k] += ll_step[k];
  }
}

// ---------------------------------------------------------------------------
// dust2/interpolate/interpolate.hpp:14-269 and spline.hpp:10-90: piecewise
// constant, piecewise linear and natural cubic spline interpolation.  The
// interval is found by a left-to-right scan (the reference bisects); the
// arithmetic of each value follows the reference so that results agree
// bit for bit.  y[i * n_series + j]: time i, series j.
// ---------------------------------------------------------------------------
namespace interp {

enum { constant = 0, linear = 1, spline = 2 };

// Slopes of the natural cubic spline through every series
// (calculate_a, calculate_b, solve_tridiagonal).
inline std::vector<double> spline_slopes(const std::vector<double>& t,
                                         const std::vector<double>& y, size_t n_series) {
  const size_t n = t.size();
  std::vector<double> k(y.size());
  for (size_t j = 0; j < n_series; ++j) {
    std::vector<double> a0(n), a1(n), a2(n), b(n);
    a1[0] = 2 / (t[1] - t[0]);
    a2[0] = 1 / (t[1] - t[0]);
    b[0] = 3 * (y[n_series + j] - y[j]) / ((t[1] - t[0]) * (t[1] - t[0]));
    for (size_t i = 1; i < n - 1; ++i) {
      const double h0 = t[i] - t[i - 1], h1 = t[i + 1] - t[i];
      a0[i] = 1 / h0;
      a1[i] = 2 * (1 / h0 + 1 / h1);
      a2[i] = 1 / h1;
      const double d0 = y[i * n_series + j] - y[(i - 1) * n_series + j];
      const double d1 = y[(i + 1) * n_series + j] - y[i * n_series + j];
      b[i] = 3 * (d0 / (h0 * h0) + d1 / (h1 * h1));
    }
    const double hl = t[n - 1] - t[n - 2];
    a0[n - 1] = 1 / hl;
    a1[n - 1] = 2 / hl;
    b[n - 1] = 3 * (y[(n - 1) * n_series + j] - y[(n - 2) * n_series + j]) / (hl * hl);
    for (size_t i = 1; i < n; ++i) {
      const double fac = a0[i] / a1[i - 1];
      a1[i] -= fac * a2[i - 1];
      b[i] -= fac * b[i - 1];
    }
    b[n - 1] /= a1[n - 1];
    for (size_t ii = 1; ii < n; ++ii) {
      const size_t i = n - 1 - ii;
      b[i] = (b[i] - a2[i] * b[i + 1]) / a1[i];
    }
    for (size_t i = 0; i < n; ++i) k[i * n_series + j] = b[i];
  }
  return k;
}

// 0: ok, 1: before the first time, 2: after the last time.
inline int eval(int kind, const std::vector<double>& t, const std::vector<double>& y,
                const std::vector<double>& k, size_t n_series, double z, double* out) {
  const size_t n = t.size();
  if (!(z >= t[0])) return 1;
  if (z > t[n - 1] && kind != constant) return 2;
  size_t i = 0;
  while (i + 1 < n && t[i + 1] <= z) ++i;
  for (size_t j = 0; j < n_series; ++j) {
    const double y0 = y[i * n_series + j];
    if (kind == constant || i == n - 1) {
      out[j] = y0;
      continue;
    }
    const double t0 = t[i], t1 = t[i + 1], y1 = y[(i + 1) * n_series + j];
    if (kind == linear) {
      out[j] = n_series == 1 ? y0 + (y1 - y0) * (z - t0) / (t1 - t0)
                             : y0 + (y1 - y0) * ((z - t0) / (t1 - t0));
      continue;
    }
    const double k0 = k[i * n_series + j], k1 = k[(i + 1) * n_series + j];
    const double r = (z - t0) / (t1 - t0);
    const double a = k0 * (t1 - t0) - (y1 - y0);
    const double b = -k1 * (t1 - t0) + (y1 - y0);
    out[j] = (1 - r) * y0 + r * y1 + r * (1 - r) * (a * (1 - r) + b * r);
  }
  return 0;
}

}  // namespace interp

// ---------------------------------------------------------------------------
// The SIR of inst/examples/sir.cpp with beta interpolated in time, the form
// odin2 generates for `beta <- interpolate(beta_t, beta_y, kind)`: update
// evaluates beta at the step's start time (dust2::interpolate, its errors
// thrown as the reference's messages).  Parameter row: N, I0, (unused),
// gamma, exp_noise, kind, n_knots, beta_t[64], beta_y[64].
// ---------------------------------------------------------------------------
struct sir_tv {
  static constexpr size_t max_knots = 64;
  static constexpr size_t n_state = 5;
  static constexpr size_t n_pars = 7 + 2 * max_knots;
  struct shared_state {
    double N, I0, gamma, exp_noise;
    int kind;
    std::vector<double> t, y, k;
  };
  using data_type = sir::data_type;
  static shared_state build_shared(const double* p) {
    shared_state sh;
    sh.N = p[0];
    sh.I0 = p[1];
    sh.gamma = p[3];
    sh.exp_noise = p[4];
    sh.kind = static_cast<int>(p[5]);
    const size_t n = static_cast<size_t>(p[6]);
    sh.t.assign(p + 7, p + 7 + n);
    sh.y.assign(p + 7 + max_knots, p + 7 + max_knots + n);
    sh.k = sh.kind == interp::spline ? interp::spline_slopes(sh.t, sh.y, 1)
                                     : std::vector<double>(n, 0.0);
    return sh;
  }
  static constexpr size_t n_data = 1;
  static data_type build_data(const double* d) { return sir::build_data(d); }
  static constexpr double zero_period = sir::zero_period;
  static constexpr size_t zero_first = sir::zero_first;
  static constexpr size_t zero_count = sir::zero_count;

  static double beta(double time, const shared_state& sh) {
    double b = 0;
    const int rc = interp::eval(sh.kind, sh.t, sh.y, sh.k, 1, time, &b);
    if (rc != 0) {
      std::ostringstream msg;
      if (rc == 1) {
        msg << "Tried to interpolate at time = " << time << ", which is " << sh.t[0] - time
            << " before the first time (" << sh.t[0] << ")";
      } else {
        msg << "Tried to interpolate at time = " << time << ", which is "
            << time - sh.t.back() << " after the last time (" << sh.t.back() << ")";
      }
      throw std::runtime_error(msg.str());
    }
    return b;
  }
  static void initial(double time, const shared_state& sh, rng_state& rng, double* y) {
    sir::initial(time, sir::shared_state{sh.N, sh.I0, 0, sh.gamma, sh.exp_noise}, rng, y);
  }
  static void update(double time, double dt, const double* y, const shared_state& sh,
                     rng_state& rng, double* yn) {
    const sir::shared_state s{sh.N, sh.I0, beta(time, sh), sh.gamma, sh.exp_noise};
    sir::update(time, dt, y, s, rng, yn);
  }
  static double compare_data(double time, const double* y, const data_type& d,
                             const shared_state& sh, rng_state& rng) {
    return sir::compare_data(time, y, d, sir::shared_state{sh.N, sh.I0, 0, sh.gamma,
                                                           sh.exp_noise}, rng);
  }
};

}  // namespace oracle
