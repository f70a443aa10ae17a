// Workgroup-level log-sum-exp, prefix sum and systematic resampling of one
// parameter group's particles, with each thread owning PPT consecutive
// particles (thread t owns particles t*PPT .. t*PPT+PPT-1).
//
// Replaces dust2::details::scale_log_weights and resample_weight
// (dust2/filter_details.hpp:11-74) and the per-group loop around them
// (dust2/filter.hpp:141-160).
//
// Exactness.  The reference compares each systematic point
// uu_i = tot*u/P + i*tot/P against a running sum accumulated strictly left
// to right, and `tot` is that same sequential sum.  A parallel scan rounds
// differently, so the kernel first resamples from a parallel FP64 scan and
// checks every decision against a bound on the difference between the two
// summation orders (both sums of non-negative terms: the sequential sum of
// j terms is within (j-1)u*tot of the exact value, the scan within depth*u*tot,
// u = 2^-53).  A point closer than that bound to either neighbouring partial
// sum is ambiguous; if any point in the group is, the group is redone with
// the sequential sum computed by one lane.  Resampling indices are thereby
// identical to the reference's for identical weights, at scan speed.
//
// The group total that gives the log-likelihood increment, log(tot / P) +
// max, and the systematic points is the aligned pairwise sum of the
// weights (within the same bound of the sequential total), so a group's
// log-likelihood has the same bits whatever workgroup size and particles
// per thread resample it: the pipeline at any resample width, the
// persistent kernel at any launch shape and a server launch sized for its
// largest request agree exactly.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "engine.hpp"

namespace dust2_amd {

constexpr int kWave = 64;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  return v;
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, kWave));
  return v;
}

// Inclusive scan across the wave.
__device__ __forceinline__ double wave_scan(double v, int lane) {
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
    const double n = __shfl_up(v, o, kWave);
    if (lane >= o) v += n;
  }
  return v;
}

// Cross-wave steps: each wave writes one value to scratch[wave]; after the
// barrier every lane reads scratch[lane % n_wave], so each aligned segment
// of n_wave lanes holds all the waves' values and combines them with
// log2(n_wave) lane exchanges (n_wave = BLOCK / 64, a power of two up to
// 16) instead of a loop of n_wave LDS reads per lane.
template <int BLOCK>
__device__ __forceinline__ double block_max(double v, double* scratch) {
  constexpr int n_wave = BLOCK / kWave;
  static_assert(n_wave >= 1 && n_wave <= 16 && (n_wave & (n_wave - 1)) == 0, "block size");
  const int lane = threadIdx.x % kWave;
  const int wave = threadIdx.x / kWave;
  v = wave_max(v);
  if constexpr (n_wave == 1) return v;
  if (lane == 0) scratch[wave] = v;
  __syncthreads();
  double r = scratch[lane % n_wave];
#pragma unroll
  for (int o = n_wave / 2; o > 0; o >>= 1) r = fmax(r, __shfl_xor(r, o, kWave));
  __syncthreads();
  return r;
}

template <int BLOCK>
__device__ __forceinline__ int block_or(int v, double* scratch) {
  constexpr int n_wave = BLOCK / kWave;
  int* s = reinterpret_cast<int*>(scratch);
  const int lane = threadIdx.x % kWave;
  const int wave = threadIdx.x / kWave;
  const int any = __any(v) ? 1 : 0;
  if constexpr (n_wave == 1) return any;
  if (lane == 0) s[wave] = any;
  __syncthreads();
  const int r = __any(s[lane % n_wave]) ? 1 : 0;
  __syncthreads();
  return r;
}

// Exclusive scan of one value per thread across the block; returns the
// exclusive prefix and writes the block total to *total.  The waves'
// totals are scanned by every n_wave-lane segment (Hillis-Steele), so the
// summation depth is log2(64) + log2(n_wave) + 1.
//
// With pair_total, also the aligned pairwise sum of the threads' values
// (lane butterflies within the wave, then across the waves), in every
// lane.  When each thread's value is itself the pairwise sum of its PPT
// consecutive elements (PPT a power of two), this is the pairwise sum of
// the elements over the aligned binary tree of their indices, zero-padded:
// the same bits whatever the block size, particles per thread or padding,
// since a subtree of zeros adds exactly 0.
template <int BLOCK>
__device__ __forceinline__ double block_exclusive_scan(double v, double* scratch, double* total,
                                                       double* pair_total = nullptr) {
  const int lane = threadIdx.x % kWave;
  const int wave = threadIdx.x / kWave;
  constexpr int n_wave = BLOCK / kWave;
  static_assert(n_wave >= 1 && n_wave <= 16 && (n_wave & (n_wave - 1)) == 0, "block size");
  const double inc = wave_scan(v, lane);
  double pw = v;
  if (pair_total) {
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) pw += __shfl_xor(pw, o, kWave);
  }
  if constexpr (n_wave == 1) {
    *total = __shfl(inc, kWave - 1, kWave);
    if (pair_total) *pair_total = pw;
    return inc - v;
  }
  if (lane == kWave - 1) {
    scratch[wave] = inc;
    if (pair_total) scratch[16 + wave] = pw;
  }
  __syncthreads();
  const int sl = lane % n_wave;
  double s = scratch[sl];
#pragma unroll
  for (int o = 1; o < n_wave; o <<= 1) {
    const double n = __shfl_up(s, o, n_wave);
    if (sl >= o) s += n;
  }
  if (pair_total) {
    double ps = scratch[16 + sl];
#pragma unroll
    for (int o = 1; o < n_wave; o <<= 1) ps += __shfl_xor(ps, o, kWave);
    *pair_total = ps;
  }
  // lane k of the segment holds the sum of waves 0..k
  const double tot = __shfl(s, n_wave - 1, kWave);
  const double before = __shfl(s, wave > 0 ? wave - 1 : 0, kWave);
  const double wave_prefix = wave > 0 ? before : 0.0;
  __syncthreads();
  *total = tot;
  return wave_prefix + inc - v;
}

// Pairwise sum of N consecutive values (N a power of two) over the aligned
// binary tree of their indices.
template <int N>
__device__ __forceinline__ double pairwise_sum(const double* e) {
  if constexpr (N == 1) {
    return e[0];
  } else {
    return pairwise_sum<N / 2>(e) + pairwise_sum<N / 2>(e + N / 2);
  }
}

// First index j in [0, P) with cum[j] >= x, or P if none.
__device__ __forceinline__ int lower_bound_lds(const double* cum, int P, double x) {
  int lo = 0;
  int hi = P;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (cum[mid] < x) {
      lo = mid + 1;
    } else {
      hi = mid;
    }
  }
  return lo;
}

// Result of one group's weighting/resampling round.
struct resample_result {
  double ll_increment;  // added to the group's log-likelihood
  bool resampled;       // false: identity (all log-weights exactly 0)
  int n_fallback;       // 1 if the sequential path was taken
};

// w[k]: log-weights of this thread's particles (k < PPT, particle
// p = tid*PPT + k, valid when p < P).  On return idx[k] holds the ancestor
// of slot p (identity when not resampled).  `cum` is P doubles of LDS.
// `u_draw` is called by thread 0 only, and only when a uniform is needed
// (it advances the group's filter stream, dust2/filter.hpp:153).
template <int BLOCK, int PPT, typename UDraw>
__device__ resample_result block_resample(double (&w)[PPT], int (&idx)[PPT], int P,
                                          double* cum, double* scratch, UDraw u_draw) {
  const int tid = threadIdx.x;
  const int p0 = tid * PPT;
  resample_result res{0.0, false, 0};

  // any(w != 0) (dust2/filter.hpp:149-150); NaN counts as non-zero
  int nz = 0;
#pragma unroll
  for (int k = 0; k < PPT; ++k) {
    if (p0 + k < P && !(w[k] == 0)) nz = 1;
  }
  if (!block_or<BLOCK>(nz, scratch)) {
#pragma unroll
    for (int k = 0; k < PPT; ++k) idx[k] = p0 + k;
    return res;
  }
  res.resampled = true;

  if (P == 1) {
    // scale_log_weights returns the single weight untouched; the only
    // ancestor is particle 0.  The uniform is still drawn.
    if (tid == 0) {
      res.ll_increment = w[0];
      (void)u_draw();
    }
    idx[0] = 0;
    return res;
  }

  const double neg_inf = -__builtin_inf();
  double mx = neg_inf;
#pragma unroll
  for (int k = 0; k < PPT; ++k) {
    if (p0 + k < P) {
      if (isnan(w[k])) w[k] = neg_inf;
      mx = fmax(mx, w[k]);
    }
  }
  // the uniform is drawn here whatever the maximum (filter.hpp:153 draws it
  // for every resampled group); its slot is read after the maximum's
  // barrier, and nothing writes it again in this call
  if (tid == 0) {
    scratch[kScratchDoubles - 1] = u_draw();
  }
  mx = block_max<BLOCK>(mx, scratch);
  if constexpr (BLOCK == kWave) __syncthreads();  // block_max has no barrier for one wave
  const double u = scratch[kScratchDoubles - 1];

  if (mx == neg_inf) {
    // Every particle impossible: the reference adds -Inf and its systematic
    // points are -Inf/NaN, so every slot takes ancestor 0.
    res.ll_increment = neg_inf;
#pragma unroll
    for (int k = 0; k < PPT; ++k) idx[k] = 0;
    return res;
  }

  // exp-scaled weights, block scan for the running sums, and the group's
  // total as the aligned pairwise sum of the weights: the log-likelihood
  // and the systematic points take that total, so they do not depend on
  // the workgroup size or particles per thread (nor on the engine)
  static_assert((PPT & (PPT - 1)) == 0, "particles per thread");
  double e[PPT];
#pragma unroll
  for (int k = 0; k < PPT; ++k) {
    e[k] = (p0 + k < P) ? exp(w[k] - mx) : 0.0;
  }
  const double local = pairwise_sum<PPT>(e);
  double scan_tot, tot;
  const double base = block_exclusive_scan<BLOCK>(local, scratch, &scan_tot, &tot);
  {
    double run = base;
#pragma unroll
    for (int k = 0; k < PPT; ++k) {
      run += e[k];
      if (p0 + k < P) cum[p0 + k] = run;
    }
  }
  __syncthreads();

  const double Pd = static_cast<double>(P);
  // bound on |sequential - scan| partial sums plus the effect on the
  // systematic points (see header)
  const double delta = tot * (4.0 * Pd + 256.0) * 0x1.0p-53;
  {
    const double uu0 = tot * u / Pd;
    const double du = tot / Pd;
    int amb = 0;
    int j_prev = 0;
#pragma unroll
    for (int k = 0; k < PPT; ++k) {
      const int i = p0 + k;
      if (i < P) {
        const double uu = uu0 + static_cast<double>(i) * du;
        // the thread's points are increasing, so each later point's
        // ancestor is at or after the previous one's: a few steps forward
        // from there, then a binary search over what is left
        int j = j_prev;
        if (k > 0) {
          for (int s = 0; s < 4 && j < P && cum[j] < uu; ++s) ++j;
          if (j < P && cum[j] < uu) j += lower_bound_lds(cum + j, P - j, uu);
        } else {
          j = lower_bound_lds(cum, P, uu);
        }
        j_prev = j;
        if ((j < P && cum[j] - uu <= delta) || (j > 0 && uu - cum[j - 1] <= delta)) {
          amb = 1;
        }
        idx[k] = j < P ? j : P - 1;
      }
    }
    amb = block_or<BLOCK>(amb, scratch);
    if (!amb) {
      res.ll_increment = log(tot / Pd) + mx;
      return res;
    }
  }

  // Sequential fallback: exact left-to-right running sums for the points;
  // the log-likelihood keeps the pairwise total.
  res.n_fallback = 1;
  __syncthreads();
#pragma unroll
  for (int k = 0; k < PPT; ++k) {
    if (p0 + k < P) cum[p0 + k] = e[k];
  }
  __syncthreads();
  if (tid == 0) {
    double run = 0;
    for (int j = 0; j < P; ++j) {
      run += cum[j];
      cum[j] = run;
    }
    scratch[kScratchDoubles - 2] = run;
  }
  __syncthreads();
  const double tot_seq = scratch[kScratchDoubles - 2];
  const double uu0 = tot_seq * u / Pd;
  const double du = tot_seq / Pd;
#pragma unroll
  for (int k = 0; k < PPT; ++k) {
    const int i = p0 + k;
    if (i < P) {
      const double uu = uu0 + static_cast<double>(i) * du;
      const int j = lower_bound_lds(cum, P, uu);
      idx[k] = j < P ? j : P - 1;
    }
  }
  res.ll_increment = log(tot / Pd) + mx;
  __syncthreads();
  return res;
}

}  // namespace dust2_amd
