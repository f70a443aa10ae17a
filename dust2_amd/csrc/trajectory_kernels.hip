// Trajectory export following the particle lineage, and the random choice
// of one particle per group.
//
// Replaces trajectories<T>::export_state with reorder = true
// (dust2/trajectories.hpp:128-189, reorder_state_group_ / reorder_state_single_
// at :313-343) and filter<T>::select_random_particle (dust2/filter.hpp:130-138).
//
// The filter saves, at every data time t, the post-resampling state
// [t][S][P] of each group and the ancestor index used by that resampling
// [t][P].  The exported trajectory of final particle k is the state of k at
// the last time, then of ancestor(k) at the time before, and so on back to
// the first time: one thread per exported particle walks the ancestry
// backwards, gathering one state vector per time.
#include <hip/hip_runtime.h>

#include "device_random.hpp"
#include "engine.hpp"

namespace dust2_amd {

namespace {

// traj: real [G][T][S][P]; ancestry: [G][T][P] (null: no resampling ever);
// groups: the n_groups_out groups to export; select: [G] chosen particle
// per group, or null for every particle; t_slot: [T] output slot of each
// time or -1 to skip it (null: slot t for time t, the trajectories);
// states: the S_out states to export (null: S_out = S, every state).
// out: double [slots][n_groups_out][P_out][S_out], P_out = select ? 1 : P.
template <typename real>
__global__ void __launch_bounds__(256)
    lineage_kernel(const real* __restrict__ traj, const uint32_t* __restrict__ ancestry,
                   const uint32_t* __restrict__ groups, int n_groups_out, int T, int S, int P,
                   const uint32_t* __restrict__ select, const int32_t* __restrict__ t_slot,
                   const uint32_t* __restrict__ states, int S_out, double* __restrict__ out) {
  const int p_out = select ? 1 : P;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_groups_out * p_out) return;
  const int j = i / p_out;          // output group
  const int k = i - j * p_out;      // output particle
  const size_t g = groups[j];
  uint32_t idx = select ? select[g] : static_cast<uint32_t>(k);
  const real* tg = traj + g * static_cast<size_t>(T) * S * P;
  const uint32_t* ag = ancestry ? ancestry + g * static_cast<size_t>(T) * P : nullptr;
  for (int t = T - 1; t >= 0; --t) {
    const int slot = t_slot ? t_slot[t] : t;
    if (slot >= 0) {
      const real* ts = tg + static_cast<size_t>(t) * S * P;
      double* o = out + ((static_cast<size_t>(slot) * n_groups_out + j) * p_out + k) * S_out;
      for (int c = 0; c < S_out; ++c) {
        const int src = states ? static_cast<int>(states[c]) : c;
        o[c] = static_cast<double>(ts[static_cast<size_t>(src) * P + idx]);
      }
    }
    if (ag) idx = ag[static_cast<size_t>(t) * P + idx];
  }
}

// One draw from filter stream 2g + 1 (column 1 of the group's [4][P + 2]
// stream block) per listed group: particle floor(u * P).
// Group g's second filter stream: word w at rng[g * group_words + first +
// w * word_stride].
__global__ void select_particle_kernel(uint64_t* __restrict__ rng, const uint32_t* groups,
                                       int n, int P, size_t group_words, size_t first,
                                       size_t word_stride, uint32_t* __restrict__ select) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const size_t g = groups[j];
  const size_t stride = word_stride;
  uint64_t* r = rng + g * group_words + first;
  rng256 s{r[0], r[stride], r[2 * stride], r[3 * stride]};
  const double u = random_real<double>(s);
  r[0] = s.s0;
  r[stride] = s.s1;
  r[2 * stride] = s.s2;
  r[3 * stride] = s.s3;
  select[g] = static_cast<uint32_t>(floor(u * P));
}

}  // namespace

hipError_t launch_lineage(const void* traj, const uint32_t* ancestry, const uint32_t* groups,
                          int n_groups_out, int T, int S, int P, const uint32_t* select,
                          const int32_t* t_slot, const uint32_t* states, int S_out, bool fp64,
                          double* out, hipStream_t stream) {
  const int n = n_groups_out * (select ? 1 : P);
  if (!states) S_out = S;
  const int block = 256;
  const dim3 grid((n + block - 1) / block);
  if (fp64) {
    hipLaunchKernelGGL(lineage_kernel<double>, grid, dim3(block), 0, stream,
                       static_cast<const double*>(traj), ancestry, groups, n_groups_out, T, S,
                       P, select, t_slot, states, S_out, out);
  } else {
    hipLaunchKernelGGL(lineage_kernel<float>, grid, dim3(block), 0, stream,
                       static_cast<const float*>(traj), ancestry, groups, n_groups_out, T, S,
                       P, select, t_slot, states, S_out, out);
  }
  return hipGetLastError();
}

hipError_t launch_select_particle(uint64_t* rng, const uint32_t* groups, int n, int P,
                                  size_t group_words, size_t first, size_t word_stride,
                                  uint32_t* select, hipStream_t stream) {
  const int block = 64;
  hipLaunchKernelGGL(select_particle_kernel, dim3((n + block - 1) / block), dim3(block), 0,
                     stream, rng, groups, n, P, group_words, first, word_stride, select);
  return hipGetLastError();
}

}  // namespace dust2_amd
