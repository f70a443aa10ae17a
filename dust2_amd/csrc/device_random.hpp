// Device-side random number generation and samplers for the particle
// engine (gfx950).  One xoshiro256+ stream per particle, held in registers
// as four 64-bit words while a particle is being advanced.
//
// Replaces monty::random::{generator<double>, random_real, binomial,
// exponential_rate, normal, poisson} and monty::density::poisson as called
// from the reference models (inst/examples/sir.cpp:63-66,115-119;
// dust2/filter.hpp:153).  Algorithms: xoshiro256+ (Blackman & Vigna),
// Hormann (1993) BTRS for large-mean binomials, inversion by sequential
// search for small means, inversion for the exponential, Box-Muller for the
// normal, Hormann (1993) PTRS for large-mean Poisson.
//
// Every function is written for both double and float `real`.  The double
// instantiation follows the arithmetic order of the FP64 CPU path term by
// term (the library is compiled with -ffp-contract=off), so FP64 draws are
// reproducible draw-for-draw up to last-place differences between device
// and host transcendental functions.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "engine.hpp"

namespace dust2_amd {

struct rng256 {
  uint64_t s0, s1, s2, s3;
};

__host__ __device__ __forceinline__ uint64_t rotl64(uint64_t x, int k) {
  return (x << k) | (x >> (64 - k));
}

// One xoshiro256+ step.  The update written on 32-bit halves with every
// new word a three-input XOR of old words (one v_bitop3_b32 per half, truth
// table 0x96) and the rotation by 45 as two funnel shifts: 12 vector
// instructions per step instead of 15.
//   s2' = s2 ^ s0 ^ (s1 << 17)    s1' = s1 ^ s2 ^ s0
//   s0' = s0 ^ s3 ^ s1            s3' = rotl(s3 ^ s1, 45)
__device__ __forceinline__ uint32_t xor3_u32(uint32_t a, uint32_t b, uint32_t c) {
  return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
}
__device__ __forceinline__ uint64_t join_u32(uint32_t hi, uint32_t lo) {
  return (static_cast<uint64_t>(hi) << 32) | lo;
}
__device__ __forceinline__ uint64_t rng_next(rng256& r) {
  const uint64_t out = r.s0 + r.s3;
  const uint64_t t = r.s1 << 17;
  const uint32_t l0 = static_cast<uint32_t>(r.s0), h0 = static_cast<uint32_t>(r.s0 >> 32);
  const uint32_t l1 = static_cast<uint32_t>(r.s1), h1 = static_cast<uint32_t>(r.s1 >> 32);
  const uint32_t l2 = static_cast<uint32_t>(r.s2), h2 = static_cast<uint32_t>(r.s2 >> 32);
  const uint32_t l3 = static_cast<uint32_t>(r.s3), h3 = static_cast<uint32_t>(r.s3 >> 32);
  const uint32_t lt = static_cast<uint32_t>(t), ht = static_cast<uint32_t>(t >> 32);
  r.s2 = join_u32(xor3_u32(h2, h0, ht), xor3_u32(l2, l0, lt));
  r.s1 = join_u32(xor3_u32(h1, h2, h0), xor3_u32(l1, l2, l0));
  r.s0 = join_u32(xor3_u32(h0, h3, h1), xor3_u32(l0, l3, l1));
  const uint32_t l = l3 ^ l1, h = h3 ^ h1;
  // rotl by 45 = rotr by 19: low word (l >> 19) | (h << 13), high word
  // (h >> 19) | (l << 13)
  r.s3 = join_u32(__builtin_amdgcn_alignbit(l, h, 19), __builtin_amdgcn_alignbit(h, l, 19));
  return out;
}

template <typename real>
__device__ __forceinline__ real random_real(rng256& r);

// 53 high bits -> [0, 1) on the 2^-53 grid.
template <>
__device__ __forceinline__ double random_real<double>(rng256& r) {
  return static_cast<double>(rng_next(r) >> 11) * 0x1.0p-53;
}

// 24 high bits -> [0, 1) on the 2^-24 grid.
template <>
__device__ __forceinline__ float random_real<float>(rng256& r) {
  // bits 40..63 kept in place in the high word (exact in float): the same
  // value as (x >> 40) * 2^-24, but a 32-bit conversion, where the shifted
  // form folds into a 64-bit to float conversion sequence
  return static_cast<float>(static_cast<uint32_t>(rng_next(r) >> 32) & 0xffffff00u) *
         0x1.0p-32f;
}

__device__ __forceinline__ double d_log(double x) { return log(x); }
__device__ __forceinline__ float d_log(float x) { return logf(x); }
__device__ __forceinline__ double d_exp(double x) { return exp(x); }
__device__ __forceinline__ float d_exp(float x) { return expf(x); }
__device__ __forceinline__ double d_sqrt(double x) { return sqrt(x); }
__device__ __forceinline__ float d_sqrt(float x) { return sqrtf(x); }
__device__ __forceinline__ double d_floor(double x) { return floor(x); }
__device__ __forceinline__ float d_floor(float x) { return floorf(x); }
__device__ __forceinline__ double d_fabs(double x) { return fabs(x); }
__device__ __forceinline__ float d_fabs(float x) { return fabsf(x); }
// 1 - exp(x) for x <= 0: the probability of at least one event at rate -x.
// FP64 keeps the reference's expression (inst/examples/sir.cpp:63) so draws
// reproduce.  FP32 avoids the cancellation of 1 - expf(x) at the small
// rates typical of a step: the Taylor series through x^6 for |x| < 1/16
// (truncation below 1e-11 relative), the hardware exp2 above, where
// 1 - exp(x) >= 0.06 keeps its relative accuracy to about 1e-6.
__device__ __forceinline__ double prob_from_rate(double x) { return 1 - exp(x); }
__device__ __forceinline__ float prob_from_rate(float x) {
  const float series =
      -x * (1.0f + x * (0.5f + x * (0.16666667f + x * (0.041666668f +
                                                      x * (0.008333334f + x * 0.0013888889f)))));
  const float direct = 1.0f - __builtin_amdgcn_exp2f(x * 1.44269504f);
  return x > -0.0625f ? series : direct;
}

// (1 - p)^n: FP64 as the reference's pow(q, n); FP32 through log1p so a
// p near the float resolution of 1 - p keeps its accuracy.
__device__ __forceinline__ double pow_one_minus(double q, double p, double n) {
  (void)p;
  return pow(q, n);
}
__device__ __forceinline__ float pow_one_minus(float q, float p, float n) {
  (void)q;
  return expf(n * log1pf(-p));
}

// Remainder of Stirling's series, log(k!) - [(k+1/2) log(k+1) - (k+1) +
// log(2 pi)/2]: exact values for k <= 9 (computed from the definition to 60
// digits), asymptotic series above.
template <typename real>
__device__ __forceinline__ real stirling_tail(real k) {
  if (k <= 9) {
    const int i = static_cast<int>(k);
    // a switch keeps the table in scalar constants rather than scratch
    switch (i) {
      case 0: return static_cast<real>(0.08106146679532726);
      case 1: return static_cast<real>(0.0413406959554093);
      case 2: return static_cast<real>(0.02767792568499834);
      case 3: return static_cast<real>(0.020790672103765093);
      case 4: return static_cast<real>(0.016644691189821193);
      case 5: return static_cast<real>(0.013876128823070748);
      case 6: return static_cast<real>(0.01189670994589177);
      case 7: return static_cast<real>(0.010411265261972096);
      case 8: return static_cast<real>(0.009255462182712733);
      default: return static_cast<real>(0.00833056343336287);
    }
  }
  const real x = k + 1;
  const real x2 = x * x;
  return (static_cast<real>(1.0 / 12) -
          (static_cast<real>(1.0 / 360) - static_cast<real>(1.0 / 1260) / x2) / x2) / x;
}

// Inversion by sequential search, n*p < 10, p <= 1/2:
// pmf(k) = pmf(k-1) * (p/q) * (n+1-k)/k.  Restarts with a fresh uniform if
// rounding lets the search run past n.  One exit test per step: a search
// that reaches k = n + 1 also updates f once more, a value the restart
// discards, so the draws are those of testing k > n before the update.
template <typename real>
__device__ real binomial_inversion(rng256& rng, real n, real p) {
  const real q = 1 - p;
  const real r = p / q;
  const real g = r * (n + 1);
  const real f0 = pow_one_minus(q, p, n);
  while (true) {
    real u = random_real<real>(rng);
    real f = f0;
    real k = 0;
    while (u >= f && k <= n) {
      u -= f;
      k += 1;
      f *= (g / k - r);
    }
    if (k <= n) {
      return k;
    }
  }
}

// BTRS (Hormann 1993), p <= 1/2, n*p >= 10.  The constants needed only on
// the rejection path (log(p/q), the mode term h, alpha) are computed the
// first time a proposal falls outside the squeeze, and that path is always
// evaluated in FP64: its log-factorial differences cancel terms of size
// n*log(n), beyond float resolution for the population sizes modelled.
template <typename real>
__device__ real binomial_btrs(rng256& rng, real n, real p) {
  const real q = 1 - p;
  const real spq = d_sqrt(n * p * q);
  const real b = static_cast<real>(1.15) + static_cast<real>(2.53) * spq;
  const real a = static_cast<real>(-0.0873) + static_cast<real>(0.0248) * b +
                 static_cast<real>(0.01) * p;
  const real c = n * p + static_cast<real>(0.5);
  const real v_r = static_cast<real>(0.92) - static_cast<real>(4.2) / b;
  bool have_bound = false;
  double alpha = 0, lpq = 0, m = 0, h = 0;
  const double nd = n;
  // one exit test per proposal (a single loop exit keeps the divergent
  // control flow cheap); every proposal draws u and v once, as before
  real k;
  bool accept;
  do {
    const real u = random_real<real>(rng) - static_cast<real>(0.5);
    const real v = random_real<real>(rng);
    const real us = static_cast<real>(0.5) - d_fabs(u);
    k = d_floor((2 * a / us + b) * u + c);
    accept = us >= static_cast<real>(0.07) && v <= v_r;
    if (!accept && k >= 0 && k <= n) {
      if (!have_bound) {
        const double pd = p;
        alpha = (2.83 + 5.1 / static_cast<double>(b)) * static_cast<double>(spq);
        lpq = log(pd / static_cast<double>(q));
        m = floor((nd + 1) * pd);
        h = (m + 0.5) * log(m + 1) + stirling_tail(m) + (nd - m + 0.5) * log(nd - m + 1) +
            stirling_tail(nd - m);
        have_bound = true;
      }
      const double kd = k;
      const double usd = us;
      const double lv = log(static_cast<double>(v) * alpha /
                            (static_cast<double>(a) / (usd * usd) + static_cast<double>(b)));
      const double bound = h - (kd + 0.5) * log(kd + 1) - stirling_tail(kd) -
                           (nd - kd + 0.5) * log(nd - kd + 1) - stirling_tail(nd - kd) +
                           (kd - m) * lpq;
      accept = lv <= bound;
    }
  } while (!accept);
  return k;
}

// ---- FP32 fast path ----------------------------------------------------------
// The FP32 samplers have no reference draw sequence to reproduce, so they
// use the hardware reciprocal/log/exp (1-2 ulp) and a rejection bound that
// is free of cancellation, keeping the whole draw in FP32 registers.
// Natural log from the hardware base-2 log (v_log_f32) times ln 2 in one
// rounding: every argument below is a normal number or zero.
__device__ __forceinline__ float f_rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float f_log(float x) {
  return __builtin_amdgcn_logf(x) * 0.693147181f;
}

// log(1 + x) accurate for small |x| from the hardware log and no
// reciprocal.  With u = 1 + x rounded and d = u - 1 (exact for u in
// [1/2, 2^24)), log(1 + x) = log(u) + log(1 + (x - d) / u), where
// |x - d| / u <= 2^-24, so the second term is (x - d) / u to float
// precision.  For |x| < 1/2, 1 / u is taken as 1 - x (error x^2 / u; where
// u rounds to 1 the result is x - x^2, log1p(x) to within x^3).  For
// x <= -1/2, 1 + x is exact and the term is zero; for x >= 1 it is
// dropped, an absolute error below 2^-24 against log(u) >= log 2.
__device__ __forceinline__ float f_log1p(float x) {
  const float u = 1.0f + x;
  const float d = u - 1.0f;
  return __builtin_fmaf(__builtin_amdgcn_logf(u), 0.693147181f,
                        (x - d) * fmaxf(1.0f - x, 0.0f));
}

// Two uniforms on the 2^-24 grid from one 64-bit output: bits 40..63 and
// 16..39 (the low bits of xoshiro256+ are the weak ones and are unused).
__device__ __forceinline__ void random_pair_f32(rng256& r, float& u, float& v) {
  const uint64_t x = rng_next(r);
  u = static_cast<float>(static_cast<uint32_t>(x >> 32) & 0xffffff00u) * 0x1.0p-32f;
  v = static_cast<float>(static_cast<uint32_t>(x >> 16) & 0xffffffu) * 0x1.0p-24f;
}

// Stirling-series remainder in FP32, log(k!) - (k + 1/2) log(k + 1) + (k + 1)
// - log(2 pi) / 2, for a pair of arguments a and b given r_a = 1/(a + 1)
// and r_b = 1/(b + 1) (the callers form both from one reciprocal of the
// product): the series 1/12 - 1/360 x^-2 + 1/1260 x^-4 over x = k + 1
// (truncation below 8e-9 from k = 4), exact values for k < 4, selected
// without a per-lane branch and only when some lane of the wavefront needs
// one (BTRS rarely evaluates the tail below k = 4).
__device__ __forceinline__ float stirling_series_f32(float r) {
  const float r2 = r * r;
  return (0.083333333f - (0.0027777778f - r2 * 0.00079365079f) * r2) * r;
}
__device__ __forceinline__ float stirling_small_f32(float k, float t) {
  t = k == 3 ? 0.020790672103765093f : t;
  t = k == 2 ? 0.02767792568499834f : t;
  t = k == 1 ? 0.0413406959554093f : t;
  t = k == 0 ? 0.08106146679532726f : t;
  return t;
}
__device__ __forceinline__ float stirling_tail_pair_f32(float a, float b, float ra, float rb) {
  float ta = stirling_series_f32(ra);
  float tb = stirling_series_f32(rb);
  if (__any(a < 4 || b < 4)) {
    ta = stirling_small_f32(a, ta);
    tb = stirling_small_f32(b, tb);
  }
  return ta + tb;
}

// 1/j for j = 0..115 (entry 0 unused), correctly rounded.
struct inv_int_table {
  float v[116];
  constexpr inv_int_table() : v() {
    for (int j = 1; j < 116; ++j) v[j] = 1.0f / static_cast<float>(j);
  }
};
__constant__ constexpr inv_int_table kInvInt{};

// Inversion, n*p < 10, p <= 1/2.
//
// The search restarts with a fresh uniform once k passes min(n, 110).  With
// mean below 10 the float pmf underflows to zero before k = 100, after
// which `u >= f` can no longer fail: a uniform left above the float sum of
// the pmf (probability about 1e-7 per draw, from rounding) would otherwise
// walk all the way to k = n before restarting, up to 1e6 iterations for one
// lane that the rest of its wave waits on.  Restarting at 110 instead gives
// the same draw sequence.
//
// j as a float for j = 0..115, read by scalar load beside 1/j.
struct int_float_table {
  float v[116];
  constexpr int_float_table() : v() {
    for (int j = 0; j < 116; ++j) v[j] = static_cast<float>(j);
  }
};
__constant__ constexpr int_float_table kIntF{};

// The search is uniform across the wavefront: every lane steps j together,
// so the loop needs no per-lane exit bookkeeping (which made the divergent
// form scalar-bound: about 14 scalar instructions per step against 8
// vector ones).  A lane whose search has ended keeps stepping its u and f
// (bounded: f stays a pmf value, or a value of magnitude at most r <= 1
// times one past k = n) but no longer moves its k, which is all it
// returns: five vector instructions per step.  Four steps per trip, the
// wavefront leaving when no lane is searching; 1/j and j come from
// constant tables, uniform across the wavefront.  Each lane's draw is that
// of the per-lane search.  A lane searching at the start of a trip has
// k = j - 1 <= 110, so the trip reads entries up to j + 3 <= 114.
__device__ inline float binomial_inversion_f32(rng256& rng, float n, float p) {
  const float q = 1 - p;
  const float r = p * f_rcp(q);
  const float g = r * (n + 1);
  const float f0 = __expf(n * f_log1p(-p));
  const float k_max = fminf(n, 110.0f);
  while (true) {
    float u = random_real<float>(rng);
    float f = f0;
    float k = 0;
    bool act = u >= f;
    for (int j = 1; __any(act); j += 4) {
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const float inv_j = kInvInt.v[j + i];
        const float jf = kIntF.v[j + i];
        u = u - f;
        f = f * __builtin_fmaf(g, inv_j, -r);
        k = act ? jf : k;
        act = act && u >= f && jf <= k_max;
      }
    }
    if (k <= k_max) {
      return k;
    }
  }
}

// BTRS with the acceptance test written as
//   log pmf(k) - log pmf(m) = d*log(p(n-m+1) / (q(m+1)))
//                             - (k+1/2) log1p(d/(m+1)) - (n-k+1/2) log1p(-d/(n-m+1))
//                             + tail(m) + tail(n-m) - tail(k) - tail(n-k),  d = k - m,
// the same quantity as the reference form (Stirling's formula for the
// factorials) with the n*log(n)-sized terms cancelled analytically.
struct btrs_f32 {
  float n, p;                        // trials and success probability (p <= 1/2)
  float b, a, c, two_a, v_r, alpha;  // envelope and squeeze
  float m, rm, rnm, L, tail_m;       // mode terms of the acceptance test
};

// Setup: three quarter-rate instructions (the square root, one reciprocal
// of b q (m + 1) (n - m + 1) that gives 1/b, 1/(m + 1), 1/(n - m + 1) and
// 1/(q (m + 1)), and the log of L).
__device__ __forceinline__ btrs_f32 btrs_setup_f32(float n, float p) {
  btrs_f32 s;
  const float q = 1 - p;
  const float spq = __builtin_amdgcn_sqrtf(n * p * q);
  s.n = n;
  s.p = p;
  s.b = 1.15f + 2.53f * spq;
  s.a = -0.0873f + 0.0248f * s.b + 0.01f * p;
  s.c = n * p + 0.5f;
  s.m = floorf((n + 1) * p);
  const float m1 = s.m + 1;
  const float nm1 = n - s.m + 1;
  const float qmn = q * m1 * nm1;
  const float r = f_rcp(s.b * qmn);  // 1 / (b q (m + 1) (n - m + 1))
  const float rb = qmn * r;
  const float qr = q * (s.b * r);    // 1 / ((m + 1) (n - m + 1))
  s.v_r = 0.92f - 4.2f * rb;
  s.two_a = 2 * s.a;
  s.alpha = (2.83f + 5.1f * rb) * spq;
  s.rm = nm1 * qr;
  s.rnm = m1 * qr;
  s.L = f_log(p * nm1 * (nm1 * (s.b * r)));
  s.tail_m = stirling_tail_pair_f32(s.m, n - s.m, s.rm, s.rnm);
  return s;
}

// One BTRS proposal: draws one (u, v) pair; true with the draw in k when
// accepted (by the squeeze or by the full test).  The full test compares
// v alpha us^2 <= (a + b us^2) exp(ratio) (the same as log(v alpha /
// (a / us^2 + b)) <= ratio, with a + b us^2 > 0) and forms both Stirling
// tails from one reciprocal: four quarter-rate instructions.
__device__ __forceinline__ bool btrs_try_f32(const btrs_f32& s, rng256& rng, float& k) {
  float u, v;
  random_pair_f32(rng, u, v);
  u -= 0.5f;
  const float us = 0.5f - fabsf(u);
  k = floorf((s.two_a * f_rcp(us) + s.b) * u + s.c);
  bool accept = us >= 0.07f && v <= s.v_r;
  if (!accept && k >= 0 && k <= s.n) {
    const float d = k - s.m;
    const float k1 = k + 1;
    const float nk1 = s.n - k + 1;
    const float r = f_rcp(k1 * nk1);
    const float ratio = d * s.L - (k + 0.5f) * f_log1p(d * s.rm) -
                        (s.n - k + 0.5f) * f_log1p(-d * s.rnm) + s.tail_m -
                        stirling_tail_pair_f32(k, s.n - k, nk1 * r, k1 * r);
    const float us2 = us * us;
    accept = v * s.alpha * us2 <= (s.a + s.b * us2) * __builtin_amdgcn_exp2f(ratio * 1.44269504f);
  }
  return accept;
}

__device__ inline float binomial_btrs_f32(rng256& rng, float n, float p) {
  const btrs_f32 s = btrs_setup_f32(n, p);
  // one exit test per proposal (cheap divergent control flow); every
  // proposal draws one (u, v) pair
  float k;
  while (!btrs_try_f32(s, rng, k)) {
  }
  return k;
}

template <typename real>
__device__ __forceinline__ real binomial_core(rng256& rng, real n, real q) {
  return (n * q >= 10) ? binomial_btrs<real>(rng, n, q) : binomial_inversion<real>(rng, n, q);
}

// Called, not inlined, from the step loops: a single out-of-line copy keeps
// the rejection loops' temporaries out of the caller's register budget.  The
// stream travels by value in and out (registers), not through a pointer,
// which would put it in scratch memory around every call.
#ifndef DUST2_BINOMIAL_LINKAGE
#define DUST2_BINOMIAL_LINKAGE __noinline__
#endif
struct binomial_draw_f32 {
  rng256 rng;
  float k;
};
__device__ DUST2_BINOMIAL_LINKAGE binomial_draw_f32 binomial_core_f32(rng256 rng, float n,
                                                                      float q) {
  binomial_draw_f32 d;
  d.k = (n * q >= 10) ? binomial_btrs_f32(rng, n, q) : binomial_inversion_f32(rng, n, q);
  d.rng = rng;
  return d;
}
template <>
__device__ __forceinline__ float binomial_core<float>(rng256& rng, float n, float q) {
  const binomial_draw_f32 d = binomial_core_f32(rng, n, q);
  rng = d.rng;
  return d.k;
}


// Binomial(n, p).  Returns ERR_BINOMIAL in *err (and 0) on invalid input,
// as the reference throws; deterministic streams return n*p.
template <typename real>
__device__ __forceinline__ real binomial(rng256& rng, real n, real p, bool deterministic,
                                         uint32_t& err) {
  if (n < 0 || p < 0 || p > 1) {
    err = ERR_BINOMIAL;
    return 0;
  }
  if (deterministic) {
    return n * p;
  }
  if (n == 0 || p == 0) {
    return 0;
  }
  if (p == 1) {
    return n;
  }
  const bool flip = p > static_cast<real>(0.5);
  const real q = flip ? 1 - p : p;
  const real k = binomial_core<real>(rng, n, q);
  return flip ? n - k : k;
}

// Exp(rate) by inversion of 1 - U (exact, in (0, 1]).
template <typename real>
__device__ __forceinline__ real exponential_rate(rng256& rng, real rate, bool deterministic) {
  if (deterministic) {
    return 1 / rate;
  }
  return -d_log(1 - random_real<real>(rng)) / rate;
}

// Box-Muller normal.
template <typename real>
__device__ __forceinline__ real normal(rng256& rng, real mean, real sd, bool deterministic) {
  if (deterministic) {
    return mean;
  }
  real u1, u2;
  do {
    u1 = random_real<real>(rng);
    u2 = random_real<real>(rng);
  } while (u1 <= static_cast<real>(sizeof(real) == 8 ? 2.2250738585072014e-308 : 1.17549435e-38));
  const real two_pi = static_cast<real>(2 * 3.14159265358979323846);
  const real z = d_sqrt(-2 * d_log(u1)) * (sizeof(real) == 8 ? static_cast<real>(cos(static_cast<double>(two_pi * u2)))
                                                               : static_cast<real>(cosf(static_cast<float>(two_pi * u2))));
  return mean + sd * z;
}

// Poisson(lambda): multiplication below 10, PTRS above.
template <typename real>
__device__ real poisson(rng256& rng, real lambda, bool deterministic) {
  if (deterministic) {
    return lambda;
  }
  if (lambda == 0) {
    return 0;
  }
  if (lambda < 10) {
    const real l = d_exp(-lambda);
    real k = 0;
    real prod = random_real<real>(rng);
    while (prod > l) {
      k += 1;
      prod *= random_real<real>(rng);
    }
    return k;
  }
  const real slam = d_sqrt(lambda);
  const real loglam = d_log(lambda);
  const real b = static_cast<real>(0.931) + static_cast<real>(2.53) * slam;
  const real a = static_cast<real>(-0.059) + static_cast<real>(0.02483) * b;
  const real inv_alpha = static_cast<real>(1.1239) + static_cast<real>(1.1328) / (b - static_cast<real>(3.4));
  const real v_r = static_cast<real>(0.9277) - static_cast<real>(3.6224) / (b - 2);
  while (true) {
    const real u = random_real<real>(rng) - static_cast<real>(0.5);
    const real v = random_real<real>(rng);
    const real us = static_cast<real>(0.5) - d_fabs(u);
    const real k = d_floor((2 * a / us + b) * u + lambda + static_cast<real>(0.43));
    if (us >= static_cast<real>(0.07) && v <= v_r) {
      return k;
    }
    if (k < 0 || (us < static_cast<real>(0.013) && v > us)) {
      continue;
    }
    const real lg = sizeof(real) == 8 ? static_cast<real>(lgamma(static_cast<double>(k + 1)))
                                      : static_cast<real>(lgammaf(static_cast<float>(k + 1)));
    if (d_log(v) + d_log(inv_alpha) - d_log(a / (us * us) + b) <= -lambda + k * loglam - lg) {
      return k;
    }
  }
}

// log Poisson pmf in FP64 with log(x!) supplied by the caller (it depends
// on the data only and is computed once on the host).
__device__ __forceinline__ double dpois_log(double x, double lambda, double lfact_x) {
  if (x == 0 && lambda == 0) {
    return 0;
  }
  if (lambda == 0) {
    return -__builtin_inf();
  }
  return x * log(lambda) - lambda - lfact_x;
}

}  // namespace dust2_amd
