// Internal interface between the host runtime (api.cpp) and the kernel
// launchers (*.hip).  Plain C++ types only.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace dust2_amd {

// Error codes recorded per particle; the host turns them into the
// reference's messages.
enum sampler_error : uint32_t {
  ERR_NONE = 0,
  ERR_BINOMIAL = 1,        // "Invalid call to binomial with n = .., p = .., q = .."
  ERR_ODE_MAX_STEPS = 2,   // "too many steps"
  ERR_ODE_STEP_SMALL = 3,  // "step too small"
  ERR_ODE_VANISHED = 4,    // "step size vanished"
  ERR_ODE_INIT = 5,        // "Initial step size was not finite"
  ERR_INTERP_BEFORE = 6,   // "Tried to interpolate at time = .., which is .. before .."
  ERR_INTERP_AFTER = 7,    // "Tried to interpolate at time = .., which is .. after .."
  ERR_ODE_HISTORY_LOCATE = 8,  // "Failed to locate history element"
  ERR_ODE_HISTORY_SHORT = 9,   // a delayed value in a step that left the history ring
};

// Doubles of LDS scratch every filter workgroup reserves ahead of its
// P-length weight array (per-wave partials and broadcast slots).
constexpr int kScratchDoubles = 64;

enum slot_flags : uint32_t {
  SLOT_SET_INITIAL = 1u,    // start from the model's initial state
  SLOT_DETERMINISTIC = 2u,  // samplers return their means
  SLOT_SAVE_TRAJ = 4u,      // write the post-reorder state at every time
  SLOT_AUTO_MODE = 16u,     // pipeline: the slot's own interval is t + t_offset;
                            // mode and buffer parity follow from it, and
                            // intervals past n_time do nothing (serving)
};

// Per-particle error record (first error of a particle).
struct error_record {
  uint32_t code;
  uint32_t interval;
  double a;
  double b;
};

// One parameter group's work in one launch of the filter kernel.  All
// pointers are device pointers; one workgroup runs one slot.
struct filter_slot {
  double pars[8];          // model parameters (model-specific order)
  const double* times;     // absolute data times, indexed by interval
  const double* obs;       // observations, obs[t * obs_stride + i]
  const double* lfact;     // log(obs!) in the same layout
  uint32_t obs_stride;
  uint32_t n_particles;
  uint32_t t_begin;        // first interval run by this launch
  uint32_t t_end;          // one past the last
  uint32_t flags;
  uint32_t group;          // group index, used in error records
  double dt;
  void* state;             // real SoA [n_state][n_particles]
  uint64_t* rng;           // SoA [4][n_particles + 2]; streams 0, 1 = filter
  double* ll;              // log-likelihood accumulator (in/out)
  double* time;            // system time (in/out)
  uint32_t* ancestry;      // [n_time][n_particles] or null
  void* traj;              // real [n_time][n_state][n_particles] or null
  error_record* errors;    // [n_particles]
  uint32_t* error_count;   // shared counter
  uint32_t* n_fallback;    // sequential-resample fallbacks (statistics)
};

struct launch_config {
  int block;
  int ppt;
};

// Persistent filter of a discrete-time model (models.hpp ids): every slot
// advances through intervals [t_begin, t_end) with particle state held in
// registers.  Models without a persistent kernel return
// hipErrorInvalidValue.
hipError_t launch_discrete_filter(int model, bool fp64, const filter_slot* d_slots, int n_slots,
                                  int max_particles, launch_config cfg, hipStream_t stream);

// Largest particle count one persistent workgroup holds (0: the model has
// no persistent kernel).
int discrete_filter_max_particles(int model);
launch_config discrete_filter_default_config(int n_particles);
bool discrete_filter_config_valid(launch_config cfg);

// ---- interval pipeline (state in HBM between intervals) -------------------
enum pipe_mode : uint32_t {
  PIPE_INITIAL = 1u,    // start from the model's initial state
  PIPE_GATHER = 2u,     // load from the ancestor slot
  PIPE_STEP = 4u,       // run the interval's steps and the compare draw
  PIPE_TRAJ_PREV = 8u,  // save the loaded (post-reorder) state of interval t-1
};

// One parameter group's buffers for the pipeline kernels; one slot per
// group, the same table for every launch of a run.
struct pipe_slot {
  const double* pars;          // model parameter block
  const double* obs;           // obs[t * obs_stride]
  const double* lfact;
  uint32_t obs_stride;
  uint32_t n_particles;
  uint32_t flags;              // SLOT_DETERMINISTIC
  uint32_t group;
  const double* sched_time;    // system time at the start of interval t
  const uint32_t* sched_steps; // steps in interval t
  const uint64_t* sched_zero;  // bit s: zero_every applies at step s (s < 64)
  double dt;
  void* state_a;               // real [S][P], double buffered
  void* state_b;
  uint64_t* rng;               // [4][P + 2]
  double* lambda;              // [P] FP64 log-weights of the last interval
  uint32_t* ancestors;         // [P] ancestors chosen by the last resample
  double* ll;
  uint32_t* ancestry;          // [n_time][P] or null
  void* traj;                  // real [n_time][S][P] or null
  error_record* errors;        // [P]
  uint32_t* error_count;
  uint32_t* n_fallback;
  int32_t t_offset;            // SLOT_AUTO_MODE: intervals already run
  uint32_t n_time;             // SLOT_AUTO_MODE: intervals in the slot's series
  // SIR with beta interpolated in time (interp_eval at each step's start
  // time): knots beta_t[beta_n], values beta_y, spline slopes beta_k;
  // beta_n = 0 keeps the constant beta of the parameter block
  const double* beta_t;
  const double* beta_y;
  const double* beta_k;
  int32_t beta_n;
  int32_t beta_kind;
};

// Interval, buffer parity and mode of a slot for launch interval t; false
// when the slot has nothing to do in this launch.
__host__ __device__ inline bool pipe_slot_interval(const pipe_slot& sd, int& t, int& parity,
                                                   uint32_t& mode) {
  if (!(sd.flags & SLOT_AUTO_MODE)) return true;
  t += sd.t_offset;
  if (t < 0 || static_cast<uint32_t>(t) >= sd.n_time) return false;
  parity = t & 1;
  mode = PIPE_STEP | (t == 0 ? PIPE_INITIAL : PIPE_GATHER);
  return true;
}

// Particles per thread of the resample workgroup of `block` threads (1024,
// 512 or 256) for groups of up to max_particles, or 0 when that width does
// not hold them (1024 threads hold up to 16,384 particles, 512 and 256
// threads up to 8 per thread).
int pipe_resample_ppt(int max_particles, int block = 1024);
hipError_t launch_pipe_reset(const pipe_slot* d_slots, int n_slots, hipStream_t stream);
// out[2i] = slot i's log-likelihood, out[2i+1] = its error count
hipError_t launch_slot_status(const pipe_slot* d_slots, int n_slots, double* out,
                              hipStream_t stream);
// One interval of every slot's particles (gather, steps, compare) for a
// discrete-time model (models.hpp ids).
hipError_t launch_discrete_advance(int model, bool fp64, const pipe_slot* d_slots, int n_slots,
                                   int max_particles, int t, int parity, uint32_t mode,
                                   hipStream_t stream);
hipError_t launch_pipe_resample(const pipe_slot* d_slots, int n_slots, int max_particles, int t,
                                hipStream_t stream, int block = 1024);
// The advance launch over a block map: entry b = slot << kBlockBits | block
// (the block-th run of discrete_advance_block(model) particles of that
// slot), kNoBlock for an unused entry; n_blocks entries, one workgroup
// each, every slot in SLOT_AUTO_MODE (its own interval, parity and mode).
constexpr uint32_t kBlockBits = 12;
constexpr uint32_t kNoBlock = 0xffffffffu;
int discrete_advance_block(int model, bool fp64);
hipError_t launch_discrete_advance_mapped(int model, bool fp64, const pipe_slot* d_slots,
                                          const uint32_t* d_block_map, int n_blocks, int t,
                                          hipStream_t stream);

// ---- systematic resampling of groups above one workgroup ----------------------
// (resample_kernels.hip)  One group per task; every pointer is device
// memory.  w: the P log-weights; idx: the P ancestors written; ancestry:
// [n_time][P] with row t_row written, or null; ll: += the log-likelihood
// increment, or null; frng: the group's first filter stream (words at
// frng[k * frng_stride]) the uniform is drawn from, or null to use u;
// n_fallback: += 1 per sequential fallback, or null; out3: [increment,
// resampled, fallback] or null; scratch: mb_resample_scratch_doubles(P)
// doubles, fully written before it is read.
constexpr int kMbMaxBlocks = 4096;  // at most 4096 * 8192 particles per group
struct mb_resample {
  const double* w;
  uint32_t* idx;
  uint32_t* ancestry;
  double* ll;
  uint64_t* frng;
  uint32_t frng_stride;
  double u;
  uint32_t* n_fallback;
  double* out3;
  double* scratch;
  int P;
};
int mb_resample_blocks(int P);
size_t mb_resample_scratch_doubles(int P);
hipError_t launch_mb_resample(const mb_resample* d_tasks, int n_tasks, int max_particles,
                              int t_row, hipStream_t stream);
// One task passed by value (no task table in device memory).
hipError_t launch_mb_resample_one(const mb_resample& task, int t_row, hipStream_t stream);

// ---- ODE unfilter ----------------------------------------------------------
// dust2/continuous/control.hpp:11-24 (step_size_min already clamped to
// machine epsilon, dust2/r/continuous/control.hpp:15-17)
// log(1e-4): the floor of the old error in the step-size controller
// (dust2/continuous/solver.hpp:210-211, max(error_old, 1e-4)), correctly
// rounded; both DP5 engines keep the log of the old error and take the
// controller's err^constant / max(error_old, 1e-4)^beta as one exp.
constexpr double kLogFacMin = -9.210340371976182;

struct ode_control_dev {
  uint32_t max_steps;
  double atol, rtol, step_size_min, step_size_max;
  double factor_safe, factor_min, factor_max, beta, constant;
  // times the solver must stop at, increasing (dust2/continuous/control.hpp
  // critical_times), device memory, or null
  const double* critical_times;
  int n_critical;
};

// Solver internals of one particle of a continuous-time system kept in HBM
// between calls (dust2/continuous/solver.hpp:34-73 internals: derivative,
// last step's dense output of the reset variable, controller state and
// counters).
struct ode_internals {
  double dydt[5];
  double step_size, error, t0, h;
  double c1, c2, c3, c4, c5;
  uint32_t n_steps, n_accepted, n_rejected, pad;
};

// Step times of one particle (control debug_record_step_times): the time
// of the initialisation and of the end of every accepted step, the first
// `cap` kept; count keeps the total.
struct ode_record {
  double* times;
  uint32_t* count;
  uint32_t cap;
};

struct ode_launch {
  const double* pars;        // [G][4]: N, I0, beta, gamma
  const uint8_t* active;     // [G] groups to run, or null for all
  int n_particles;
  int n_total;               // G * P
  double time_start;
  const double* times;       // [n_time]
  int n_time;
  const double* obs;         // obs[t * obs_stride + g * obs_group_stride]
  const double* lfact;       // same layout
  size_t obs_stride;
  size_t obs_group_stride;   // 0 for shared data, 1 per group
  ode_control_dev control;
  void* state;               // real [S][n_total]
  double* ll;                // [n_total]
  double* stats;             // [n_total][4] or null
  error_record* errors;      // [n_total]
  // set to 1 (a plain store) by every particle that fails; zeroed by the
  // caller; may be pinned host memory
  uint32_t* error_flag;
  // optional per-group data series (serving): group g reads times, obs and
  // lfact from offset series_offset[g], series_len[g] values, one value per
  // time; null for one series shared through obs_stride/obs_group_stride
  const int64_t* series_offset;
  const int32_t* series_len;
  // with per-group series: the start time of group g's series, or null
  // for time_start everywhere
  const double* series_time_start;
  // forward sensitivity run (FP64): parameter 0 beta, 1 gamma, 2 I0, or -1
  // for none; grad[3 * particle + sens_param] receives d ll / d theta
  int sens_param;
  double* grad;
  // state at each data time, real [n_time][S][n_total], or null
  void* traj;
};

hipError_t launch_dp5_unfilter(const ode_launch& args, bool fp64, hipStream_t stream);

// ---- continuous-time system (SIR ODE, FP64) ---------------------------------
// Runs every particle from `time` through the target times in turn; the
// state after target q goes to out[q][S][n_total] when out is given.
struct ode_sys_launch {
  const double* pars;          // [G][4]: N, I0, beta, gamma
  int n_particles;
  int n_total;
  const uint8_t* reinit;       // [G]: initialise the solver first
  double time;                 // system time at the start
  const double* targets;       // [n_targets]
  int n_targets;
  ode_control_dev control;
  double* state;               // [S][n_total]
  ode_internals* internals;    // [n_total]
  double* out;                 // [n_targets][S][n_total] or null
  ode_record record;           // step times, or times = null
  error_record* errors;        // [n_total]
  uint32_t* error_count;
};
hipError_t launch_dp5_system(const ode_sys_launch& args, hipStream_t stream);

// ---- deterministic discrete SIR unfilter and adjoint -------------------------
struct sir_unfilter_launch {
  const double* pars;           // [G][8]: N, I0, beta, gamma, exp_noise, p_IR, e_IR,
                                //   exp(-gamma)
  const uint8_t* active;        // [G] groups to run, or null for all
  int n_particles;
  int n_total;                  // G * P
  int set_initial;              // 1: model initial state, 0: start from `state`
  double dt;
  int n_time;
  const double* sched_time;     // [n_time] time at the start of interval i
  const uint32_t* sched_steps;  // [n_time] steps in interval i
  const double* obs;            // obs[i * obs_stride + g * obs_group_stride]
  const double* lfact;          // same layout
  size_t obs_stride;
  size_t obs_group_stride;
  double* state;                // [S][n_total] in/out
  double* ll;                   // [n_total]
  double* hist;                 // [total steps][2][n_total] or null (no gradient)
  double* inc;                  // [n_time][n_total] or null
  double* grad;                 // [n_total][3] or null
  double* traj;                 // [n_time][S][n_total] state at each data time, or null
  error_record* errors;         // [n_total]
  uint32_t* error_count;
  // 0: the reference's hand-written adjoint (inst/examples/sir.cpp:126-198)
  // term for term; 1: the derivative of the computed likelihood
  int exact;
};

hipError_t launch_sir_unfilter(const sir_unfilter_launch& args, hipStream_t stream);

// ---- trajectories ------------------------------------------------------------
// Lineage export: traj real [G][T][S][P], ancestry [G][T][P] or null,
// groups[n_groups_out], select [G] chosen particle or null (all particles),
// t_slot [T] output slot per time or -1 (null: every time, slot t);
// states [S_out] the states exported (null: all S);
// out double [slots][n_groups_out][P_out][S_out].
hipError_t launch_lineage(const void* traj, const uint32_t* ancestry, const uint32_t* groups,
                          int n_groups_out, int T, int S, int P, const uint32_t* select,
                          const int32_t* t_slot, const uint32_t* states, int S_out, bool fp64,
                          double* out, hipStream_t stream);
// select[g] = floor(u * P), u from group g's second filter stream (stream
// 2g + 1 of the reference's numbering), for the n listed groups; word w of
// that stream at rng[g * group_words + first + w * word_stride] (the
// [G][4][P + 2] blocks of the discrete filter: 4 (P + 2), 1, P + 2; the
// [G][2][4] filter streams of the mixed-time filter: 8, 4, 1).
hipError_t launch_select_particle(uint64_t* rng, const uint32_t* groups, int n, int P,
                                  size_t group_words, size_t first, size_t word_stride,
                                  uint32_t* select, hipStream_t stream);

// out[q * n_series + j]: series j of the interpolator (kind, t[n], y, k) at
// z[q]; err[q]: 0 or the interp_error of z[q].
hipError_t launch_interpolate(int kind, const double* t, const double* y, const double* k, int n,
                              int n_series, const double* z, int n_z, double* out, int* err,
                              hipStream_t stream);

// ---- system (state resident in HBM between calls) -------------------------
// state: real [G][S][P]; rng: [G][4][P]; kpars: [G][kstride] doubles (the
// model's kernel parameter block).
struct discrete_run_args {
  void* state;
  uint64_t* rng;
  const double* kpars;
  int kstride;
  int G, P;
  double time, dt;
  int n_steps;
  uint32_t flags;            // SLOT_SET_INITIAL, SLOT_DETERMINISTIC
  error_record* errors;      // [G][P]
  uint32_t* error_count;
};
hipError_t launch_discrete_run(int model, bool fp64, const discrete_run_args& a,
                               hipStream_t stream);
// out[g * P + p] = compare_data of each particle; obs/lfact offset by
// g * obs_group_stride
hipError_t launch_discrete_compare(int model, bool fp64, const discrete_run_args& a,
                                   const double* obs, const double* lfact, int obs_group_stride,
                                   double* out, hipStream_t stream);
hipError_t launch_reorder(const void* state_in, void* state_out, const uint32_t* idx, int G,
                          int S, int P, bool fp64, hipStream_t stream);

// ---- RNG -------------------------------------------------------------------
// Stream n = jump^n(s0) for n < n_streams, using the column-major GF(2)
// matrices J_i = jump^(2^i), i < n_mats; s0 is passed by value.  Written
// AoS (4 words per stream) when soa_stride is 0, else SoA: word w of
// stream n at out[w * soa_stride + n].
struct rng_words {
  uint64_t w[4];
};
hipError_t launch_seed_streams(const uint64_t s0[4], const uint64_t* d_mats, int n_mats,
                               uint64_t n_streams, uint64_t* d_out, uint64_t soa_stride,
                               hipStream_t stream);
// AoS raw streams <-> per-group SoA [G][4][stride] at column offset `first`.
hipError_t launch_rng_aos_to_soa(const uint64_t* aos, uint64_t* soa, int G, int per_group,
                                 int stride, hipStream_t stream);
hipError_t launch_rng_soa_to_aos(const uint64_t* soa, uint64_t* aos, int G, int per_group,
                                 int stride, hipStream_t stream);

// ---- vector-ALU calibration (calibration_kernels.hip) --------------------------
// n_blocks workgroups of four wavefronts, each wavefront issuing
// iters * valu_probe_fmas_per_iteration(fp64) independent fused
// multiply-adds (plus loop overhead); out: n_blocks * 256 doubles.
int valu_probe_fmas_per_iteration(bool fp64);
hipError_t launch_valu_probe(bool fp64, int n_blocks, int iters, double* out,
                             hipStream_t stream);

// ---- test entry points -----------------------------------------------------
hipError_t launch_resample_test(const double* w, int n, double u, uint32_t* idx,
                                double* ll_increment, hipStream_t stream);
hipError_t launch_sample_test(int kind, bool fp64, uint64_t* rng_aos, int n_streams,
                              int n_draws, const double* a, const double* b, int det,
                              double* out, uint32_t* err, hipStream_t stream);

}  // namespace dust2_amd
