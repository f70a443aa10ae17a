// Systematic resampling of one group across many workgroups, for groups
// too large for one workgroup's LDS (above 16,384 particles).  The same
// arithmetic and exactness rule as block_resample.hpp (which replaces
// dust2::details::scale_log_weights and resample_weight,
// dust2/filter_details.hpp:11-74): the maximum, the exp-scaled weights, a
// parallel FP64 prefix sum, the systematic points tot*u/P + i*tot/P located
// by binary search, every decision checked against the bound
// tot*(4P+256)*2^-53 on the difference between the parallel and the
// reference's left-to-right summation; a group with any ambiguous point is
// redone with the sequential sum.
//
// Five launches per round, each over (blocks of the group x groups):
//   partials  per-block maximum and any-nonzero flag
//   scale     global maximum, the uniform (drawn once per group from its
//             filter stream), per-block sums of exp(w - max)
//   cum       canonical block offsets (a sequential sum of the block sums,
//             the same in every block) and the inclusive prefix sum of the
//             group, written to the scratch array cum[P]
//   search    each point's ancestor by binary search in cum; ambiguity flag
//             per block
//   finish    one workgroup per group: the log-likelihood increment, or the
//             sequential fallback
// Scratch per group (doubles): mb_header + 4 * blocks + P, laid out as
// header | block max | block nonzero | block sum | block ambiguous | cum.
#include <hip/hip_runtime.h>

#include "block_resample.hpp"
#include "device_random.hpp"
#include "engine.hpp"

namespace dust2_amd {

namespace {

constexpr int kMbBlock = 1024;
constexpr int kMbPpt = 8;
constexpr int kMbChunk = kMbBlock * kMbPpt;
constexpr int kMbHeader = 8;  // [0] max, [1] nonzero, [2] u, [3] tot

__device__ __forceinline__ int mb_blocks(int P) { return (P + kMbChunk - 1) / kMbChunk; }

struct mb_view {
  double* hdr;
  double* part_max;
  double* part_nz;
  double* part_sum;
  double* part_amb;
  double* cum;
};

__device__ __forceinline__ mb_view mb_layout(const mb_resample& t) {
  const int B = mb_blocks(t.P);
  double* s = t.scratch;
  return mb_view{s, s + kMbHeader, s + kMbHeader + B, s + kMbHeader + 2 * B,
                 s + kMbHeader + 3 * B, s + kMbHeader + 4 * B};
}

// NaN log-weights count as impossible (scale_log_weights, filter_details.hpp:18-25)
__device__ __forceinline__ double scaled(double w, double mx) {
  return isnan(w) ? 0.0 : exp(w - mx);
}

__global__ void __launch_bounds__(kMbBlock) mb_partials_kernel(const mb_resample* tasks,
                                                                mb_resample one) {
  __shared__ double sm[kScratchDoubles];
  const mb_resample t = tasks ? tasks[blockIdx.y] : one;
  const int b = blockIdx.x;
  if (b >= mb_blocks(t.P)) return;
  const mb_view v = mb_layout(t);
  const int p0 = b * kMbChunk + threadIdx.x * kMbPpt;
  double mx = -__builtin_inf();
  int nz = 0;
#pragma unroll
  for (int k = 0; k < kMbPpt; ++k) {
    const int p = p0 + k;
    if (p < t.P) {
      const double w = t.w[p];
      if (!(w == 0)) nz = 1;  // any(w != 0), NaN included (dust2/filter.hpp:149-150)
      if (!isnan(w)) mx = fmax(mx, w);
    }
  }
  mx = block_max<kMbBlock>(mx, sm);
  nz = block_or<kMbBlock>(nz, sm);
  if (threadIdx.x == 0) {
    v.part_max[b] = mx;
    v.part_nz[b] = nz;
  }
}

__global__ void __launch_bounds__(kMbBlock) mb_scale_kernel(const mb_resample* tasks,
                                                             mb_resample one) {
  __shared__ double sm[kScratchDoubles];
  const mb_resample t = tasks ? tasks[blockIdx.y] : one;
  const int b = blockIdx.x;
  const int B = mb_blocks(t.P);
  if (b >= B) return;
  const mb_view v = mb_layout(t);
  double mx = -__builtin_inf();
  int nz = 0;
  for (int i = threadIdx.x; i < B; i += kMbBlock) {
    mx = fmax(mx, v.part_max[i]);
    if (v.part_nz[i] != 0) nz = 1;
  }
  mx = block_max<kMbBlock>(mx, sm);
  nz = block_or<kMbBlock>(nz, sm);
  if (b == 0 && threadIdx.x == 0) {
    double u = 0;
    if (nz) {
      // one uniform per group from its filter stream (dust2/filter.hpp:153),
      // drawn only when some weight is non-zero
      if (t.frng) {
        const uint32_t st = t.frng_stride;
        rng256 r{t.frng[0], t.frng[st], t.frng[2 * st], t.frng[3 * st]};
        u = random_real<double>(r);
        t.frng[0] = r.s0;
        t.frng[st] = r.s1;
        t.frng[2 * st] = r.s2;
        t.frng[3 * st] = r.s3;
      } else {
        u = t.u;
      }
    }
    v.hdr[0] = mx;
    v.hdr[1] = nz;
    v.hdr[2] = u;
  }
  if (!nz || mx == -__builtin_inf()) return;
  const int p0 = b * kMbChunk + threadIdx.x * kMbPpt;
  double local = 0;
#pragma unroll
  for (int k = 0; k < kMbPpt; ++k) {
    if (p0 + k < t.P) local += scaled(t.w[p0 + k], mx);
  }
  double total;
  (void)block_exclusive_scan<kMbBlock>(local, sm, &total);
  if (threadIdx.x == 0) v.part_sum[b] = total;
}

__global__ void __launch_bounds__(kMbBlock) mb_cum_kernel(const mb_resample* tasks,
                                                           mb_resample one) {
  __shared__ double sm[kScratchDoubles];
  __shared__ double sums[kMbMaxBlocks];
  const mb_resample t = tasks ? tasks[blockIdx.y] : one;
  const int b = blockIdx.x;
  const int B = mb_blocks(t.P);
  if (b >= B) return;
  const mb_view v = mb_layout(t);
  const double mx = v.hdr[0];
  if (v.hdr[1] == 0 || mx == -__builtin_inf()) return;
  for (int i = threadIdx.x; i < B; i += kMbBlock) sums[i] = v.part_sum[i];
  __syncthreads();
  if (threadIdx.x == 0) {
    // block offsets and the total by one sequential sum over the block
    // sums, identical in every block
    double run = 0, base = 0;
    for (int i = 0; i < B; ++i) {
      if (i == b) base = run;
      run += sums[i];
    }
    sm[kScratchDoubles - 1] = base;
    sm[kScratchDoubles - 2] = run;
  }
  __syncthreads();
  const double base = sm[kScratchDoubles - 1];
  const double tot = sm[kScratchDoubles - 2];
  __syncthreads();
  const int p0 = b * kMbChunk + threadIdx.x * kMbPpt;
  double e[kMbPpt];
  double local = 0;
#pragma unroll
  for (int k = 0; k < kMbPpt; ++k) {
    e[k] = p0 + k < t.P ? scaled(t.w[p0 + k], mx) : 0.0;
    local += e[k];
  }
  double total;
  double run = base + block_exclusive_scan<kMbBlock>(local, sm, &total);
#pragma unroll
  for (int k = 0; k < kMbPpt; ++k) {
    run += e[k];
    if (p0 + k < t.P) v.cum[p0 + k] = run;
  }
  if (b == 0 && threadIdx.x == 0) v.hdr[3] = tot;
}

// First index j in [0, P) with cum[j] >= x, or P if none.
__device__ __forceinline__ int lower_bound_global(const double* cum, int P, double x) {
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

__global__ void __launch_bounds__(kMbBlock) mb_search_kernel(const mb_resample* tasks,
                                                              mb_resample one, int t_row) {
  __shared__ double sm[kScratchDoubles];
  const mb_resample t = tasks ? tasks[blockIdx.y] : one;
  const int b = blockIdx.x;
  if (b >= mb_blocks(t.P)) return;
  const mb_view v = mb_layout(t);
  const int P = t.P;
  const int p0 = b * kMbChunk + threadIdx.x * kMbPpt;
  uint32_t* anc = t.ancestry ? t.ancestry + static_cast<size_t>(t_row) * P : nullptr;
  const double mx = v.hdr[0];
  const bool nz = v.hdr[1] != 0;
  int amb = 0;
  if (!nz || mx == -__builtin_inf()) {
    // identity when every log-weight is zero; ancestor 0 for every slot
    // when every particle is impossible (as block_resample)
#pragma unroll
    for (int k = 0; k < kMbPpt; ++k) {
      const int i = p0 + k;
      if (i < P) {
        const uint32_t j = nz ? 0u : static_cast<uint32_t>(i);
        t.idx[i] = j;
        if (anc) anc[i] = j;
      }
    }
  } else {
    const double tot = v.hdr[3];
    const double u = v.hdr[2];
    const double Pd = static_cast<double>(P);
    const double uu0 = tot * u / Pd;
    const double du = tot / Pd;
    const double delta = tot * (4.0 * Pd + 256.0) * 0x1.0p-53;
#pragma unroll
    for (int k = 0; k < kMbPpt; ++k) {
      const int i = p0 + k;
      if (i < P) {
        const double uu = uu0 + static_cast<double>(i) * du;
        const int j = lower_bound_global(v.cum, P, uu);
        if ((j < P && v.cum[j] - uu <= delta) || (j > 0 && uu - v.cum[j - 1] <= delta)) {
          amb = 1;
        }
        const uint32_t a = static_cast<uint32_t>(j < P ? j : P - 1);
        t.idx[i] = a;
        if (anc) anc[i] = a;
      }
    }
  }
  amb = block_or<kMbBlock>(amb, sm);
  if (threadIdx.x == 0) v.part_amb[b] = amb;
}

__global__ void __launch_bounds__(kMbBlock) mb_finish_kernel(const mb_resample* tasks,
                                                              mb_resample one, int t_row) {
  __shared__ double sm[kScratchDoubles];
  const mb_resample t = tasks ? tasks[blockIdx.x] : one;
  const mb_view v = mb_layout(t);
  const int P = t.P;
  const int B = mb_blocks(P);
  const double mx = v.hdr[0];
  const bool nz = v.hdr[1] != 0;
  const int tid = threadIdx.x;
  if (!nz) {
    if (tid == 0 && t.out3) {
      t.out3[0] = 0;
      t.out3[1] = 0;
      t.out3[2] = 0;
    }
    return;
  }
  const double Pd = static_cast<double>(P);
  if (mx == -__builtin_inf()) {
    if (tid == 0) {
      if (t.ll) t.ll[0] += -__builtin_inf();
      if (t.out3) {
        t.out3[0] = -__builtin_inf();
        t.out3[1] = 1;
        t.out3[2] = 0;
      }
    }
    return;
  }
  int amb = 0;
  for (int i = tid; i < B; i += kMbBlock) {
    if (v.part_amb[i] != 0) amb = 1;
  }
  amb = block_or<kMbBlock>(amb, sm);
  double tot = v.hdr[3];
  if (amb) {
    // sequential fallback: exact left-to-right running sums
    for (int j = tid; j < P; j += kMbBlock) v.cum[j] = scaled(t.w[j], mx);
    __syncthreads();
    if (tid == 0) {
      double run = 0;
      for (int j = 0; j < P; ++j) {
        run += v.cum[j];
        v.cum[j] = run;
      }
      sm[kScratchDoubles - 1] = run;
    }
    __syncthreads();
    tot = sm[kScratchDoubles - 1];
    const double u = v.hdr[2];
    const double uu0 = tot * u / Pd;
    const double du = tot / Pd;
    uint32_t* anc = t.ancestry ? t.ancestry + static_cast<size_t>(t_row) * P : nullptr;
    for (int i = tid; i < P; i += kMbBlock) {
      const double uu = uu0 + static_cast<double>(i) * du;
      const int j = lower_bound_global(v.cum, P, uu);
      const uint32_t a = static_cast<uint32_t>(j < P ? j : P - 1);
      t.idx[i] = a;
      if (anc) anc[i] = a;
    }
  }
  if (tid == 0) {
    const double inc = log(tot / Pd) + mx;
    if (t.ll) t.ll[0] += inc;
    if (amb && t.n_fallback) atomicAdd(t.n_fallback, 1u);
    if (t.out3) {
      t.out3[0] = inc;
      t.out3[1] = 1;
      t.out3[2] = amb ? 1 : 0;
    }
  }
}

}  // namespace

int mb_resample_blocks(int P) { return (P + kMbChunk - 1) / kMbChunk; }

size_t mb_resample_scratch_doubles(int P) {
  return static_cast<size_t>(kMbHeader) + 4 * static_cast<size_t>(mb_resample_blocks(P)) +
         static_cast<size_t>(P);
}

hipError_t launch_mb_resample(const mb_resample* d_tasks, int n_tasks, int max_particles,
                              int t_row, hipStream_t stream) {
  if (n_tasks == 0) return hipSuccess;
  const int B = mb_resample_blocks(max_particles);
  if (B > kMbMaxBlocks) return hipErrorInvalidValue;
  const dim3 grid(B, n_tasks);
  const mb_resample none{};
  hipLaunchKernelGGL(mb_partials_kernel, grid, dim3(kMbBlock), 0, stream, d_tasks, none);
  hipLaunchKernelGGL(mb_scale_kernel, grid, dim3(kMbBlock), 0, stream, d_tasks, none);
  hipLaunchKernelGGL(mb_cum_kernel, grid, dim3(kMbBlock), 0, stream, d_tasks, none);
  hipLaunchKernelGGL(mb_search_kernel, grid, dim3(kMbBlock), 0, stream, d_tasks, none, t_row);
  hipLaunchKernelGGL(mb_finish_kernel, dim3(n_tasks), dim3(kMbBlock), 0, stream, d_tasks, none,
                     t_row);
  return hipGetLastError();
}

hipError_t launch_mb_resample_one(const mb_resample& task, int t_row, hipStream_t stream) {
  const int B = mb_resample_blocks(task.P);
  if (task.P < 1 || B > kMbMaxBlocks) return hipErrorInvalidValue;
  const dim3 grid(B, 1);
  const mb_resample* none = nullptr;
  hipLaunchKernelGGL(mb_partials_kernel, grid, dim3(kMbBlock), 0, stream, none, task);
  hipLaunchKernelGGL(mb_scale_kernel, grid, dim3(kMbBlock), 0, stream, none, task);
  hipLaunchKernelGGL(mb_cum_kernel, grid, dim3(kMbBlock), 0, stream, none, task);
  hipLaunchKernelGGL(mb_search_kernel, grid, dim3(kMbBlock), 0, stream, none, task, t_row);
  hipLaunchKernelGGL(mb_finish_kernel, dim3(1), dim3(kMbBlock), 0, stream, none, task, t_row);
  return hipGetLastError();
}

}  // namespace dust2_amd
