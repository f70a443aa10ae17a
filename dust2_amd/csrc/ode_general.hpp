// Host/device interface of the general continuous-time engine
// (ode_general_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "engine.hpp"

namespace dust2_amd {

// Word offsets of one particle's solver internals (FP64 words): the
// derivative, the last step's dense output (history_step), the controller
// state and counters, the time of the last root of each event, the history
// ring's count and initial values, and the number of events found
// (dust2/continuous/solver.hpp:34-73 internals).
struct gen_layout {
  int dydt, t0, t1, h, c1, c2, c3, c4, c5, step_size, error, n_steps, n_accepted, n_rejected,
      last_event, hist_count, hist_first_t0, hist_init, event_count, words;
};

__host__ __device__ constexpr gen_layout make_gen_layout(int nv, int ny, int me, int nh) {
  gen_layout L{};
  int o = 0;
  L.dydt = o;
  o += nv;
  L.t0 = o++;
  L.t1 = o++;
  L.h = o++;
  L.c1 = o;
  o += ny;
  L.c2 = o;
  o += nv;
  L.c3 = o;
  o += nv;
  L.c4 = o;
  o += nv;
  L.c5 = o;
  o += nv;
  L.step_size = o++;
  L.error = o++;
  L.n_steps = o++;
  L.n_accepted = o++;
  L.n_rejected = o++;
  L.last_event = o;
  o += me;
  L.hist_count = o++;
  L.hist_first_t0 = o++;
  L.hist_init = o;
  o += nh;
  L.event_count = o++;
  L.words = o;
  return L;
}

struct gen_launch {
  int model;                 // odem::model_id
  const double* pars;        // [G][n_pars]
  int n_pars;
  int n_particles;
  int n_total;
  const uint8_t* reinit;     // [G]: initialise the solver first
  const uint8_t* active;     // [G]: groups to run, or null for all
  double time;               // system time at the start
  const double* targets;     // [n_targets]
  int n_targets;
  ode_control_dev control;
  double* state;             // [S][n_total], S = n_var + n_special + n_output
  double* internals;         // [n_total][words]
  int words;
  double* out;               // [n_targets][S][n_total] or null
  ode_record record;         // step times, or times = null
  double* events;            // [n_total][event_cap][3] or null
  uint32_t event_cap;
  double* history;           // [n_total][history_cap][3 + 5 NH] or null
  uint32_t history_cap;
  int save_history;          // keep the history ring (always with delays)
  double dt;                 // mixed-time update interval, 0 for none
  const uint64_t* rng;       // [n_total][4] system streams (mixed-time models)
  int deterministic;         // updates take the mean of their draws
  error_record* errors;      // [n_total]
  uint32_t* error_count;
};

struct gen_compare_launch {
  int model;
  const double* pars;        // [G][n_pars]
  int n_pars;
  int n_particles;
  int n_total;
  double time;
  const double* state;       // [S][n_total]
  const double* data;        // [G or 1][n_data]
  int shared_data;
  double* out;               // [n_total]
  int accumulate;            // add to out instead of writing it
  const uint8_t* active;     // [G] or null for all
};

struct gen_resample_launch {
  const double* ll_step;     // [G][P] log-weights
  uint64_t* filter_rng;      // [G][2][4]: streams 2g (resampling) and 2g + 1
  double* ll;                // [G] accumulated log-likelihood
  uint32_t* idx;             // [G][P] ancestors
  const uint32_t* groups;    // [n_groups] groups to resample (one workgroup each)
  int n_groups;
  int n_particles;           // at most 8192
  uint32_t* fallback_count;
};

struct gen_reorder_launch {
  const uint32_t* idx;       // [G][P] ancestors within the group
  const double* state_in;    // [S][n_total]
  double* state_out;
  int S;
  const double* words_in;    // [n_total][words] solver internals
  double* words_out;
  int words;
  int n_particles;
  int n_total;
  double* traj;              // [G][n_time][S][P] or null
  uint32_t* ancestry;        // [G][n_time][P] or null
  int n_time;
  int t;
  const uint8_t* active;     // [G] or null: inactive groups keep their order
};

int gen_words(int model);
int gen_history_words(int model);
hipError_t launch_general_system(const gen_launch& args, hipStream_t stream);
hipError_t launch_general_compare(const gen_compare_launch& args, hipStream_t stream);
hipError_t launch_general_resample(const gen_resample_launch& args, hipStream_t stream);
hipError_t launch_general_reorder(const gen_reorder_launch& args, hipStream_t stream);

}  // namespace dust2_amd
