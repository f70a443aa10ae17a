/* dust2_amd: MI355X-native batched particle engine behind dust2's entry
 * points.
 *
 * C ABI only: plain pointers, sizes and status codes.  Every function that
 * can fail returns 0 on success and non-zero on failure, writing a
 * NUL-terminated message (the reference's wording where one exists) into
 * `err` (capacity `errlen`).  Handles own their device memory; they are not
 * thread-safe individually, distinct handles may be driven concurrently.
 *
 * Conventions shared with the reference's R/cpp11 layer:
 *   - parameters: one row of model parameters per group (pars[g * n_pars]);
 *     sir:    N, I0, beta, gamma, exp_noise   (inst/examples/sir.cpp:18-24)
 *     sirode: N, I0, beta, gamma              (inst/examples/sirode.cpp:15-20)
 *   - data: data[t][g][j], n_groups_data == 1 means data shared by all
 *     groups (dust2/discrete/system.hpp:267,273); NaN marks missing data.
 *   - state in/out: [group][particle][state] (the memory order of the
 *     reference's [state, particle, group] arrays, dust2/r/system.hpp:44-56).
 *   - indices are 0-based (the C++ side of the reference,
 *     dust2/r/helpers.hpp:279-306).
 *   - RNG state: 4 x uint64 per stream (xoshiro256+), [stream][word]; filter
 *     layout per group is [2 filter streams | P particle streams]
 *     (dust2/r/filter.hpp:182-238).
 *   - seed: when `seed_state` is non-NULL its first `n_seed_streams` streams
 *     are used verbatim and later streams continue by jumping; otherwise the
 *     integer `seed` seeds stream 0 and stream k is stream 0 jumped k times
 *     (vignettes/design.Rmd:72-75).
 */
#ifndef DUST2_AMD_H
#define DUST2_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DUST2_FP64 0
#define DUST2_FP32 1

/* Filter engines (results are identical; only speed differs):
 *   PERSISTENT: one launch per run, each group's particles held in one
 *               workgroup's registers for every interval;
 *   PIPELINE:   per interval an advance launch (one thread per particle,
 *               state in HBM) and a resample launch (one workgroup per
 *               group), optionally replayed from a captured HIP graph. */
#define DUST2_ENGINE_AUTO 0
#define DUST2_ENGINE_PERSISTENT 1
#define DUST2_ENGINE_PIPELINE 2

typedef struct dust2_filter dust2_filter;
typedef struct dust2_system dust2_system;
typedef struct dust2_unfilter dust2_unfilter;
typedef struct dust2_unfilter_sir dust2_unfilter_sir;
typedef struct dust2_ode_system dust2_ode_system;
typedef struct dust2_ode_general dust2_ode_general;
typedef struct dust2_filter_ode dust2_filter_ode;
typedef struct dust2_unfilter_ode dust2_unfilter_ode;

/* ---- library ------------------------------------------------------------ */

/* 0 if a usable GPU is present and the kernels load; error text otherwise. */
int dust2_amd_init(char* err, size_t errlen);
const char* dust2_amd_version(void);

/* ---- bootstrap particle filter (stochastic SIR) ---------------------------
 * Replaces dust2_discrete_filter_alloc<sir> (dust2/r/discrete/filter.hpp:12-88)
 * registered as dust2_filter_sir_alloc (inst/template/discrete/compare.cpp:1-9).
 * `time` must be strictly increasing and after time_start
 * (dust2/r/helpers.hpp:266-276); dt > 0 and integer or 1/integer
 * (dust2/r/helpers.hpp:232-262).  `precision` is DUST2_FP64 (FP64 state and
 * samplers) or DUST2_FP32 (FP32 state and samplers, FP64 log-weights,
 * resampling sums and log-likelihood). */
int dust2_filter_sir_alloc(const double* pars, size_t n_groups, size_t n_particles,
                           double time_start, const double* time, size_t n_time, double dt,
                           const double* data, size_t n_groups_data,
                           const uint64_t* seed_state, size_t n_seed_streams, uint64_t seed,
                           int precision, dust2_filter** out, char* err, size_t errlen);

/* Age-structured stochastic SEIR with 16 age groups (no reference model; the
 * same generated-class contract).  Parameter row: beta, sigma, gamma,
 * exp_noise, I0, N[16], C[16*16] (row-major contact matrix); data: 16
 * weekly incidence values per group per time; state: S[16] E[16] I[16]
 * R[16] inc[16].  Runs on the pipeline engine: DUST2_FP32 with the force of
 * infection on the matrix cores, DUST2_FP64 in the CPU's arithmetic order.
 * All dust2_filter_* methods apply. */
int dust2_filter_seir16_alloc(const double* pars, size_t n_groups, size_t n_particles,
                              double time_start, const double* time, size_t n_time, double dt,
                              const double* data, size_t n_groups_data,
                              const uint64_t* seed_state, size_t n_seed_streams, uint64_t seed,
                              int precision, dust2_filter** out, char* err, size_t errlen);

/* The filter of any discrete-time model by name ("sir", "seir16"; a model
 * without compare_data, such as "walk", is refused), as the reference
 * instantiates one filter<dust_discrete<T>> per generated class T
 * (dust2/r/discrete/filter.hpp:12-88, inst/template/discrete/compare.cpp). */
int dust2_filter_discrete_alloc(const char* model, const double* pars, size_t n_groups,
                                size_t n_particles, double time_start, const double* time,
                                size_t n_time, double dt, const double* data,
                                size_t n_groups_data, const uint64_t* seed_state,
                                size_t n_seed_streams, uint64_t seed, int precision,
                                dust2_filter** out, char* err, size_t errlen);

/* Replaces dust2_filter_run<T> (dust2/r/filter.hpp:21-52) / filter<T>::run
 * (dust2/filter.hpp:42-104), with the argument checks of dust_likelihood_run
 * (R/interface-likelihood.R:83-98); indices are 0-based.
 *   initial: NULL (model initial state) or [G][P][S].
 *   save_trajectories: keep the states of index_state at every data time.
 *   save_snapshots: NULL or n_snapshots data times, strictly increasing, at
 *     which the whole state is kept ("All elements of 'save_snapshots' must
 *     be found in 'time'").
 *   index_state: NULL (every state) or n_index_state distinct states kept
 *     in the trajectories (trajectories<T>::set_index_and_reset,
 *     dust2/trajectories.hpp:31-45); snapshots keep every state.
 *   index_group: NULL (all groups) or n_index distinct group indices.
 * ll_out receives one log-likelihood per selected group, in index order. */
int dust2_filter_run(dust2_filter* h, const double* initial, int save_trajectories,
                     const double* save_snapshots, size_t n_snapshots,
                     const size_t* index_state, size_t n_index_state,
                     const size_t* index_group, size_t n_index, double* ll_out,
                     char* err, size_t errlen);

/* Replaces dust2_filter_update_pars<T> (dust2/r/filter.hpp:10-19); sir
 * updates I0, beta and gamma (inst/examples/sir.cpp:89-93); N and
 * exp_noise are constant.  pars: n_index rows of n_pars. */
int dust2_filter_update_pars(dust2_filter* h, const double* pars,
                                 const size_t* index_group, size_t n_index,
                                 char* err, size_t errlen);

/* Replaces dust2_filter_last_state (dust2/r/filter.hpp): out [G][P][S]. */
int dust2_filter_last_state(dust2_filter* h, double* out, char* err, size_t errlen);

/* Replaces dust2_filter_last_trajectories (dust2/r/filter.hpp:54-95) with
 * trajectories<T>::export_state(reorder = true) (dust2/trajectories.hpp:128-189):
 * each exported particle's trajectory follows its ancestry back from the
 * last time.  select_random_particle != 0 exports one particle per group,
 * chosen once per run from filter stream 2g + 1 (filter<T>::select_random_particle,
 * dust2/filter.hpp:130-138).  out [n_time][n_groups of the last run][P or 1]
 * [n_index_state of the last run, or S].  Requires save_trajectories on the
 * last run ("Trajectories are not current"). */
int dust2_filter_last_trajectories(dust2_filter* h, int select_random_particle, double* out,
                                   char* err, size_t errlen);

/* Replaces dust2_filter_last_snapshots (dust2/r/filter.hpp:97-139) with
 * trajectories<T>::export_snapshots(reorder = true) (dust2/trajectories.hpp:
 * 191-283): the whole state at the snapshot times along each exported
 * particle's ancestry.  out [n_snapshots][n_groups of the last run][P or 1][S];
 * "Snapshots are not current" if the last run saved none. */
int dust2_filter_last_snapshots(dust2_filter* h, int select_random_particle, double* out,
                                char* err, size_t errlen);

/* dust2_filter_last_state with select_random_particle (dust2/r/filter.hpp:
 * 141-180): the state of the chosen particle (the same choice as the
 * trajectory export) of each group of the last run: out [n_groups][S]. */
int dust2_filter_last_state_selected(dust2_filter* h, double* out, char* err, size_t errlen);

/* The saved post-resampling states as stored, without following the
 * ancestry: out [n_time][G][P][S].  No counterpart in the reference. */
int dust2_filter_last_trajectories_raw(dust2_filter* h, double* out, char* err,
                                       size_t errlen);

/* Ancestor indices of the last run: out [n_time][G][P] (0-based). */
int dust2_filter_last_ancestry(dust2_filter* h, uint32_t* out, char* err, size_t errlen);

/* Replaces dust2_filter_rng_state / set_rng_state (dust2/r/filter.hpp:182-238):
 * G*(P+2) streams, per group [2 filter | P particle]. */
int dust2_filter_rng_state(dust2_filter* h, uint64_t* out, char* err, size_t errlen);
int dust2_filter_set_rng_state(dust2_filter* h, const uint64_t* state, char* err,
                                   size_t errlen);

/* Launch shape of the persistent kernel (threads per workgroup, particles
 * per thread); results do not depend on it.  Chosen automatically unless
 * set; invalid shapes are rejected. */
int dust2_filter_set_launch(dust2_filter* h, int block, int particles_per_thread,
                                char* err, size_t errlen);

/* Statistics of the last run (out[7]): [0] resampling rounds that needed the
 * sequential sum, [1] GPU milliseconds of the run (device events), [2] the
 * engine used, [3] / [4] mean advance / resample kernel milliseconds per
 * interval (pipeline engine with profiling on), [5] / [6] the persistent
 * engine's threads per workgroup and particles per thread. */
int dust2_filter_last_stats(dust2_filter* h, double* out, char* err, size_t errlen);

/* SIR with beta interpolated in time: from the next run each step uses
 * beta at its start time from the interpolator (kind 0 piecewise constant,
 * 1 linear, 2 natural cubic spline) through knots t[n_t], values
 * y[n_groups_y][n_t] (n_groups_y 1: shared, or one row per group), in place
 * of the constant beta; kind < 0 removes it.  The model odin2 generates for
 * `beta <- interpolate(beta_t, beta_y, kind)` in inst/examples/sir.cpp's
 * update, with dust2::interpolate (dust2/interpolate/interpolate.hpp:69-268)
 * and its errors for a step time outside the knots, reported per particle.
 * Runs on the pipeline engine. */
int dust2_filter_set_beta_schedule(dust2_filter* h, int kind, const double* t, size_t n_t,
                                   const double* y, size_t n_groups_y, char* err,
                                   size_t errlen);

/* Number of independent launch chains (1-4, each on its own stream) the
 * pipeline engine splits the groups into, so that one chain's resampling
 * overlaps another chain's advance (default 2; dust2_filter_autotune
 * chooses it by measurement).  Results do not depend on it.  No
 * counterpart in the reference. */
int dust2_filter_set_pipeline_chains(dust2_filter* h, int chains, char* err, size_t errlen);

/* Threads per resample workgroup of the pipeline engine: 1024 (default),
 * 512 or 256, the particles per thread following from the group size
 * (1024 threads hold up to 16,384 particles, 512 and 256 threads up to 8
 * per thread; a width that does not hold the groups is rejected).
 * dust2_filter_autotune chooses it with the chain count.  Results do not
 * depend on it: the resampling indices are exact and the log-likelihood
 * takes the weights' aligned pairwise sum, whose bits do not depend on the
 * workgroup shape.  No counterpart in the reference. */
int dust2_filter_set_pipeline_resample(dust2_filter* h, int block, char* err, size_t errlen);

/* Engine selection (DUST2_ENGINE_*), HIP-graph replay of the pipeline
 * (default on) and per-kernel event profiling (default off, disables the
 * graph).  AUTO runs the pipeline, except the persistent kernel for 256 or
 * more SIR groups of at most 512 particles.  No counterpart in the
 * reference (its one engine is the CPU). */
int dust2_filter_set_engine(dust2_filter* h, int engine, int use_graph, int profile,
                                char* err, size_t errlen);

/* Autotuner (no counterpart in the reference): times a run of the groups in
 * index_group (NULL: all) with every candidate engine and launch shape -- the
 * persistent kernel at each threads x particles-per-thread shape that holds
 * the group, and the interval pipeline -- and keeps the fastest.  Results do
 * not depend on the choice; the RNG streams and state are restored, so the
 * object is left as it was apart from the engine choice (trajectories of
 * earlier runs are no longer current).  The pipeline is tried with one to
 * four launch chains, each with resample workgroups of 1024, 512 and 256
 * threads where they hold the groups.  report (may be NULL): up to
 * report_rows rows of [engine, threads per workgroup (persistent) or per
 * resample workgroup (pipeline), particles per thread (persistent) or
 * launch chains (pipeline), milliseconds]; *n_report receives the number
 * of candidates (at most 26). */
int dust2_filter_autotune(dust2_filter* h, const size_t* index_group, size_t n_index,
                          double* report, size_t report_rows, size_t* n_report, char* err,
                          size_t errlen);

void dust2_filter_free(dust2_filter* h);

/* ---- discrete-time system ----------------------------------------------------
 * Replaces dust2_discrete_alloc<T> (dust2/r/discrete/system.hpp:12-54) and
 * the per-model dust2_system_<name>_* methods (inst/template/system.cpp:1-54):
 * one set of methods for every discrete-time model, the model named at
 * allocation ("sir"; "seir16"; "walk", inst/examples/walk.cpp, parameter
 * row sd, len (1..8), random_initial (0 or 1), no data).  All groups must
 * have the same state length (dust2/r/helpers.hpp:333-369). */
int dust2_system_discrete_alloc(const char* model, const double* pars, size_t n_groups,
                                size_t n_particles, double time, double dt,
                                const uint64_t* seed_state, size_t n_seed_streams, uint64_t seed,
                                int deterministic, int precision, dust2_system** out, char* err,
                                size_t errlen);
/* dust2_system_discrete_alloc("sir", ...), registered by the reference as
 * dust2_system_sir_alloc (inst/template/system.cpp:1-4). */
int dust2_system_sir_alloc(const double* pars, size_t n_groups, size_t n_particles,
                           double time, double dt, const uint64_t* seed_state,
                           size_t n_seed_streams, uint64_t seed, int deterministic,
                           int precision, dust2_system** out, char* err, size_t errlen);
int dust2_system_discrete_run_to_time(dust2_system* h, double time, char* err, size_t errlen);
int dust2_system_discrete_simulate(dust2_system* h, const double* times, size_t n_times,
                              double* out, char* err, size_t errlen);
int dust2_system_discrete_state(dust2_system* h, double* out, char* err, size_t errlen);
int dust2_system_discrete_set_state(dust2_system* h, const double* state, char* err,
                               size_t errlen);
int dust2_system_discrete_set_state_initial(dust2_system* h, char* err, size_t errlen);
int dust2_system_discrete_time(dust2_system* h, double* out, char* err, size_t errlen);
int dust2_system_discrete_set_time(dust2_system* h, double time, char* err, size_t errlen);
int dust2_system_discrete_update_pars(dust2_system* h, const double* pars, char* err,
                                 size_t errlen);
int dust2_system_discrete_compare_data(dust2_system* h, const double* data, size_t n_groups_data,
                                  double* out, char* err, size_t errlen);
int dust2_system_discrete_reorder(dust2_system* h, const uint32_t* index, char* err,
                             size_t errlen);
int dust2_system_discrete_rng_state(dust2_system* h, uint64_t* out, char* err, size_t errlen);
int dust2_system_discrete_set_rng_state(dust2_system* h, const uint64_t* state, char* err,
                                   size_t errlen);
void dust2_system_discrete_free(dust2_system* h);

/* ---- deterministic likelihood of the SIR ODE (Dormand-Prince 5(4)) --------
 * Replaces dust2_continuous_unfilter_alloc<sirode>
 * (dust2/r/continuous/unfilter.hpp:14-58, registered as
 * dust2_unfilter_sirode_alloc, inst/template/compare.cpp:6-29).
 * control: NULL (dust_ode_control() defaults, R/interface-continuous.R:51-79)
 * or [max_steps, atol, rtol, step_size_min, step_size_max]; step_size_min is
 * clamped to machine epsilon (dust2/r/continuous/control.hpp:15-17).
 * precision: DUST2_FP64, or DUST2_FP32 for FP32 state and stages with the
 * error norm, step size and controller kept in FP64. */
int dust2_unfilter_sirode_alloc(const double* pars, size_t n_groups, size_t n_particles,
                                double time_start, const double* time, size_t n_time,
                                const double* data, size_t n_groups_data, const double* control,
                                int precision, dust2_unfilter** out, char* err, size_t errlen);

/* Replaces dust2_unfilter_run<T> (dust2/r/unfilter.hpp:21-62) /
 * unfilter<T>::run (dust2/unfilter.hpp:42-72): ll_out receives
 * n_particles log-likelihoods per selected group.  save_trajectories keeps
 * the states of index_state at every data time (trajectories_.add,
 * dust2/unfilter.hpp:66); save_snapshots, index_state and index_group as
 * dust2_filter_run. */
int dust2_unfilter_sirode_run(dust2_unfilter* h, int save_trajectories,
                              const double* save_snapshots, size_t n_snapshots,
                              const size_t* index_state, size_t n_index_state,
                              const size_t* index_group, size_t n_index, double* ll_out,
                              char* err, size_t errlen);

/* Replaces dust2_unfilter_last_trajectories (dust2/r/unfilter.hpp:64-96),
 * no reordering: out [n_time][n_groups of the last run][P][n_index_state
 * of the last run, or S]. */
int dust2_unfilter_sirode_last_trajectories(dust2_unfilter* h, double* out, char* err,
                                            size_t errlen);
/* Replaces dust2_unfilter_last_snapshots (dust2/r/unfilter.hpp:97-128,
 * registered per model as dust2_unfilter_<name>_last_snapshots,
 * inst/template/compare.cpp:22): the whole state at the snapshot times of
 * the last run, out [n_snapshots][n_groups of the last run][P][S];
 * "Snapshots are not current" if it saved none. */
int dust2_unfilter_sirode_last_snapshots(dust2_unfilter* h, double* out, char* err,
                                         size_t errlen);
int dust2_unfilter_sirode_update_pars(dust2_unfilter* h, const double* pars,
                                      const size_t* index_group, size_t n_index, char* err,
                                      size_t errlen);
/* out [G][P][S] */
int dust2_unfilter_sirode_last_state(dust2_unfilter* h, double* out, char* err, size_t errlen);
/* Solver counters of the last run, as dust_system_internals reports them
 * (dust2/r/continuous/system.hpp:56-158): out [G*P][4] = n_steps,
 * n_steps_accepted, n_steps_rejected, step_size; kernel_ms may be NULL. */
int dust2_unfilter_sirode_last_stats(dust2_unfilter* h, double* out, double* kernel_ms,
                                     char* err, size_t errlen);
/* Likelihood and its gradient with respect to (beta, gamma, I0) by forward
 * sensitivities of the ODE, integrated with the same steps as the state (the
 * sensitivities do not enter the error control), so ll_out is exactly the
 * plain run's.  No counterpart in the reference, whose only gradient is the
 * discrete adjoint (dust2/unfilter.hpp:77-229); the interface mirrors that
 * one.  FP64 only. */
int dust2_unfilter_sirode_run_gradient(dust2_unfilter* h, int save_trajectories,
                                       const double* save_snapshots, size_t n_snapshots,
                                       const size_t* index_state, size_t n_index_state,
                                       const size_t* index_group, size_t n_index,
                                       double* ll_out, char* err, size_t errlen);
/* out [n_index of the last run][P][3] = d ll / d (beta, gamma, I0). */
int dust2_unfilter_sirode_last_gradient(dust2_unfilter* h, double* out, char* err,
                                        size_t errlen);
void dust2_unfilter_sirode_free(dust2_unfilter* h);
/* control.critical_times (dust2/continuous/control.hpp:26-39) for the ODE
 * unfilter: the solver stops at each of these times (strictly increasing)
 * as solver::run does (dust2/continuous/solver.hpp:243-259). */
int dust2_unfilter_sirode_set_critical_times(dust2_unfilter* h, const double* times, size_t n,
                                             char* err, size_t errlen);

/* ---- continuous-time system (SIR ODE, FP64) --------------------------------
 * Replaces dust2_continuous_alloc<sirode> (dust2/r/continuous/system.hpp)
 * and the dust2_system_sirode_* methods (inst/template/system.cpp:1-54)
 * over dust_continuous (dust2/continuous/system.hpp:32-330): state
 * [G][P][5] as the other entry points; the solver internals of every
 * particle stay on the device between calls and are re-initialised after
 * set_state, set_state_initial, set_time and update_pars.  control: the
 * five dust_ode_control() values or NULL; critical_times[n_critical];
 * record_step_times: 0, or the number of step times kept per particle
 * (control debug_record_step_times).  Errors lock the system as the
 * reference does (dust2/r/system.hpp:9-16). */
int dust2_system_sirode_alloc(const double* pars, size_t n_groups, size_t n_particles,
                              double time, const double* control, const double* critical_times,
                              size_t n_critical, size_t record_step_times,
                              dust2_ode_system** out, char* err, size_t errlen);
int dust2_system_sirode_run_to_time(dust2_ode_system* h, double time, char* err,
                                    size_t errlen);
int dust2_system_sirode_simulate(dust2_ode_system* h, const double* times, size_t n_times,
                                 double* out, char* err, size_t errlen);
int dust2_system_sirode_state(dust2_ode_system* h, double* out, char* err, size_t errlen);
int dust2_system_sirode_set_state(dust2_ode_system* h, const double* state, char* err,
                                  size_t errlen);
int dust2_system_sirode_set_state_initial(dust2_ode_system* h, char* err, size_t errlen);
int dust2_system_sirode_time(dust2_ode_system* h, double* out, char* err, size_t errlen);
int dust2_system_sirode_set_time(dust2_ode_system* h, double time, char* err, size_t errlen);
int dust2_system_sirode_update_pars(dust2_ode_system* h, const double* pars, char* err,
                                    size_t errlen);
int dust2_system_sirode_compare_data(dust2_ode_system* h, const double* data,
                                     size_t n_groups_data, double* out, char* err,
                                     size_t errlen);
int dust2_system_sirode_reorder(dust2_ode_system* h, const uint32_t* index, char* err,
                                size_t errlen);
/* dust_system_internals (dust2/r/continuous/system.hpp, solver internals
 * dust2/continuous/solver.hpp:34-73): stats[G*P][4] = n_steps,
 * n_steps_accepted, n_steps_rejected, step_size; with recording,
 * step_times[G*P][record_step_times] and their counts n_step_times[G*P]. */
int dust2_system_sirode_internals(dust2_ode_system* h, double* stats, double* step_times,
                                  size_t* n_step_times, char* err, size_t errlen);
void dust2_system_sirode_free(dust2_ode_system* h);

/* ---- general continuous-time systems: events, carried variables, delays --
 * One set of entry points for the continuous-time models of the reference's
 * tests and examples that use events, special (carried) variables, periodic
 * resets and delays, chosen by name at alloc: "bounce"
 * (tests/testthat/examples/event.cpp), "change" (event-time.cpp),
 * "eventzero" (event-zero.cpp), "specialevent" (specialevent.cpp),
 * "sirdelay" (delay.cpp), "logdelay" (a delayed logistic equation whose
 * rhs reads the delayed state) and the mixed-time models "special"
 * (special.cpp) and "malaria" (inst/examples/malaria.cpp), which with
 * dt > 0 apply their discrete update every dt (system.hpp:544-585).  They replace the per-model
 * dust2_system_<name>_* registrations that inst/template/system.cpp:1-54
 * generates over dust_continuous (dust2/continuous/system.hpp:32-585) with
 * solver::step / run / apply_events (dust2/continuous/solver.hpp:152-262,
 * 359-435), history and delays (dust2/continuous/history.hpp,
 * delays.hpp); an R shim maps dust2_system_bounce_alloc(...) to
 * dust2_system_ode_alloc("bounce", ...).  FP64.
 *
 * sizes[9] from dust2_system_ode_model: n_var, n_special, n_output, n_pars,
 * max_events, history variables, has_delays, mixed, n_data.  State [G][P][S] with
 * S = n_var + n_special + n_output (the reference's packing order);
 * pars[G][n_pars]. */
int dust2_system_ode_model(const char* model, size_t* sizes, char* err, size_t errlen);
/* dt: 0, or for a mixed-time model the update interval (an integer or the
 * inverse of one); seed_state / seed / deterministic as
 * dust2_system_sir_alloc, for the streams of mixed-time updates.
 * control: the five dust_ode_control() values or NULL; critical_times;
 * record_step_times: step times kept per particle (0: off); save_history
 * (control save_history; always on for a model with delays) with
 * history_cap steps kept per particle, the most recent; event_cap: events
 * kept per particle (the first).  Errors lock the system as the reference
 * does (dust2/r/system.hpp:9-16). */
int dust2_system_ode_alloc(const char* model, const double* pars, size_t n_groups,
                           size_t n_particles, double time, double dt, const double* control,
                           const double* critical_times, size_t n_critical,
                           size_t record_step_times, size_t save_history, size_t history_cap,
                           size_t event_cap, const uint64_t* seed_state, size_t n_seed_streams,
                           uint64_t seed, int deterministic, dust2_ode_general** out, char* err,
                           size_t errlen);
/* the particles' random number streams ([G*P][4] words, stream k for
 * particle k) of a mixed-time system (dust2/r/system.hpp rng_state /
 * set_rng_state) */
int dust2_system_ode_rng_state(dust2_ode_general* h, uint64_t* out, char* err, size_t errlen);
int dust2_system_ode_set_rng_state(dust2_ode_general* h, const uint64_t* state, char* err,
                                   size_t errlen);
/* compare_data (dust2/continuous/system.hpp:287-312) for a model with data
 * (malaria: tested, positive): data[n_groups_data][n_data], out[G*P]. */
int dust2_system_ode_compare_data(dust2_ode_general* h, const double* data,
                                  size_t n_groups_data, double* out, char* err, size_t errlen);
int dust2_system_ode_run_to_time(dust2_ode_general* h, double time, char* err, size_t errlen);
int dust2_system_ode_simulate(dust2_ode_general* h, const double* times, size_t n_times,
                              double* out, char* err, size_t errlen);
int dust2_system_ode_state(dust2_ode_general* h, double* out, char* err, size_t errlen);
int dust2_system_ode_set_state(dust2_ode_general* h, const double* state, char* err,
                               size_t errlen);
int dust2_system_ode_set_state_initial(dust2_ode_general* h, char* err, size_t errlen);
int dust2_system_ode_time(dust2_ode_general* h, double* out, char* err, size_t errlen);
int dust2_system_ode_set_time(dust2_ode_general* h, double time, char* err, size_t errlen);
int dust2_system_ode_update_pars(dust2_ode_general* h, const double* pars, char* err,
                                 size_t errlen);
int dust2_system_ode_reorder(dust2_ode_general* h, const uint32_t* index, char* err,
                             size_t errlen);
/* dust_system_internals (dust2/r/continuous/system.hpp:56-160):
 * stats[G*P][4] = n_steps, n_steps_accepted, n_steps_rejected, step_size;
 * step_times[G*P][record_step_times] with their counts. */
int dust2_system_ode_internals(dust2_ode_general* h, double* stats, double* step_times,
                               size_t* n_step_times, char* err, size_t errlen);
/* the events found (internals$events): counts[G*P] and
 * records[G*P][event_cap][3] = time, event index (0-based), sign. */
int dust2_system_ode_events(dust2_ode_general* h, size_t* counts, double* records, char* err,
                            size_t errlen);
/* the history (internals$history): counts[G*P] of steps saved and, oldest
 * first, entries[G*P][history_cap][3 + 5 H] = t0, t1, h, then the
 * coefficients c1..c5 of the H history variables. */
int dust2_system_ode_history(dust2_ode_general* h, size_t* counts, double* entries, char* err,
                             size_t errlen);
void dust2_system_ode_free(dust2_ode_general* h);

/* ---- particle filter over a mixed-time model ------------------------------
 * dust_filter_create for a model of the general continuous-time engine with
 * a discrete update and data (malaria), replacing the filter registrations
 * of inst/template/compare.cpp:31-64 over filter<dust_continuous<T>>
 * (dust2/filter.hpp:42-175): per data time run_to_time, compare_data, the
 * log-sum-exp and systematic resampling of each group (exact indices) and
 * the reorder of state and solver internals.  Streams [2 filter | P
 * particle] per group as dust2_filter_sir_alloc (dust2/r/discrete/
 * filter.hpp:63-74).  data[n_time][n_groups_data][n_data]; dt: the update
 * interval (0: none); control: dust_ode_control() values or NULL.  FP64;
 * at most 8192 particles per group. */
int dust2_filter_ode_alloc(const char* model, const double* pars, size_t n_groups,
                           size_t n_particles, double time_start, const double* time,
                           size_t n_time, const double* data, size_t n_groups_data, double dt,
                           const double* control, const uint64_t* seed_state,
                           size_t n_seed_streams, uint64_t seed, dust2_filter_ode** out,
                           char* err, size_t errlen);
/* initial: [G][P][S] or NULL (the model's initial state); save_snapshots,
 * index_state and index_group as dust2_filter_run; ll_out[n_index] */
int dust2_filter_ode_run(dust2_filter_ode* h, const double* initial, int save_trajectories,
                         const double* save_snapshots, size_t n_snapshots,
                         const size_t* index_state, size_t n_index_state,
                         const size_t* index_group, size_t n_index, double* ll_out, char* err,
                         size_t errlen);
int dust2_filter_ode_update_pars(dust2_filter_ode* h, const double* pars, char* err,
                                 size_t errlen);
int dust2_filter_ode_last_state(dust2_filter_ode* h, double* out, char* err, size_t errlen);
/* out[n_time][groups of the last run][P or 1][n_index_state or S], each
 * particle's lineage (dust2/trajectories.hpp:128-189); select_random_particle
 * != 0 exports one particle per group, drawn once per run from the group's
 * second filter stream (filter<T>::select_random_particle,
 * dust2/filter.hpp:130-138; dust2_filter_last_trajectories,
 * dust2/r/filter.hpp:54-95) */
int dust2_filter_ode_last_trajectories(dust2_filter_ode* h, int select_random_particle,
                                       double* out, char* err, size_t errlen);
/* out[n_snapshots][groups of the last run][P or 1][S] along the same lineages
 * (dust2/r/filter.hpp:97-139, dust2/trajectories.hpp:191-283) */
int dust2_filter_ode_last_snapshots(dust2_filter_ode* h, int select_random_particle,
                                    double* out, char* err, size_t errlen);
/* dust2_filter_last_state with select_random_particle (dust2/r/filter.hpp:
 * 141-180): the chosen particle's state for each group of the last run,
 * out[n_groups of the last run][S]. */
int dust2_filter_ode_last_state_selected(dust2_filter_ode* h, double* out, char* err,
                                         size_t errlen);
/* [G][P + 2][4] words: per group the 2 filter streams, then the particles' */
int dust2_filter_ode_rng_state(dust2_filter_ode* h, uint64_t* out, char* err, size_t errlen);
int dust2_filter_ode_set_rng_state(dust2_filter_ode* h, const uint64_t* state, char* err,
                                   size_t errlen);
void dust2_filter_ode_free(dust2_filter_ode* h);

/* ---- deterministic likelihood over a general model with data ------------
 * dust_unfilter_create for a model of the general engine with compare_data
 * (malaria), replacing the unfilter registrations of
 * inst/template/compare.cpp:6-29 over unfilter<dust_continuous<T>>
 * (dust2/unfilter.hpp:42-72, 162-186): deterministic updates (their draws'
 * means, dust2/r/continuous/unfilter.hpp:38-42), per data time run_to_time
 * and compare_data, one log-likelihood per particle.  Arguments as
 * dust2_filter_ode_alloc and dust2_filter_ode_run; ll_out[n_index][P];
 * trajectories out[n_time][groups of the last run][P][n_index_state or S];
 * snapshots out[n_snapshots][groups of the last run][P][S]. */
int dust2_unfilter_ode_alloc(const char* model, const double* pars, size_t n_groups,
                             size_t n_particles, double time_start, const double* time,
                             size_t n_time, const double* data, size_t n_groups_data, double dt,
                             const double* control, dust2_unfilter_ode** out, char* err,
                             size_t errlen);
int dust2_unfilter_ode_run(dust2_unfilter_ode* h, const double* initial, int save_trajectories,
                           const double* save_snapshots, size_t n_snapshots,
                           const size_t* index_state, size_t n_index_state,
                           const size_t* index_group, size_t n_index, double* ll_out,
                           char* err, size_t errlen);
int dust2_unfilter_ode_update_pars(dust2_unfilter_ode* h, const double* pars, char* err,
                                   size_t errlen);
int dust2_unfilter_ode_last_state(dust2_unfilter_ode* h, double* out, char* err,
                                  size_t errlen);
int dust2_unfilter_ode_last_trajectories(dust2_unfilter_ode* h, double* out, char* err,
                                         size_t errlen);
int dust2_unfilter_ode_last_snapshots(dust2_unfilter_ode* h, double* out, char* err,
                                      size_t errlen);
void dust2_unfilter_ode_free(dust2_unfilter_ode* h);

/* ---- deterministic discrete-time SIR unfilter with the adjoint ----------
 * Replaces dust2_discrete_unfilter_alloc<sir> (dust2/r/discrete/unfilter.hpp:12-58),
 * registered as dust2_unfilter_sir_alloc (inst/template/discrete/compare.cpp).
 * Samplers return their means (the reference's deterministic streams); FP64.
 * Parameters as dust2_filter_sir_alloc; data[t][g] with n_groups_data 1 or G. */
int dust2_unfilter_sir_alloc(const double* pars, size_t n_groups, size_t n_particles,
                             double time_start, const double* time, size_t n_time, double dt,
                             const double* data, size_t n_groups_data,
                             dust2_unfilter_sir** out, char* err, size_t errlen);

/* Replaces dust2_unfilter_run<T> (dust2/r/unfilter.hpp:21-62) ->
 * unfilter<T>::run / run_adjoint (dust2/unfilter.hpp:42-114), arguments in
 * the same order.  initial: NULL (model initial state) or [G][P][S].
 * adjoint != 0 also computes the gradient of the log-likelihood with
 * respect to (beta, gamma, I0).  save_snapshots, index_state and
 * index_group as dust2_filter_run.  ll_out: [n_index][P], one
 * log-likelihood per particle. */
int dust2_unfilter_sir_run(dust2_unfilter_sir* h, const double* initial, int save_trajectories,
                           const double* save_snapshots, size_t n_snapshots, int adjoint,
                           const size_t* index_state, size_t n_index_state,
                           const size_t* index_group, size_t n_index, double* ll_out,
                           char* err, size_t errlen);

/* Replaces dust2_unfilter_last_trajectories (dust2/r/unfilter.hpp:64-96):
 * the states of index_state at every data time, out [n_time][n_groups of
 * the last run][P][n_index_state or S] (identical with and without the
 * adjoint). */
int dust2_unfilter_sir_last_trajectories(dust2_unfilter_sir* h, double* out, char* err,
                                         size_t errlen);
/* Replaces dust2_unfilter_last_snapshots (dust2/r/unfilter.hpp:97-128,
 * dust2_unfilter_<name>_last_snapshots of inst/template/compare.cpp:22):
 * out [n_snapshots][n_groups of the last run][P][S]. */
int dust2_unfilter_sir_last_snapshots(dust2_unfilter_sir* h, double* out, char* err,
                                      size_t errlen);

/* Replaces dust2_discrete_unfilter_last_gradient (dust2/r/unfilter.hpp:153-179):
 * out [n_index of the last run][3] = d ll / d (beta, gamma, I0).  Errors as the
 * reference when the last run had no adjoint or P > 1. */
int dust2_unfilter_sir_last_gradient(dust2_unfilter_sir* h, double* out, char* err,
                                     size_t errlen);

/* Which gradient dust2_unfilter_sir_last_gradient returns.  exact = 0 (the
 * default): the reference's hand-written adjoint term for term
 * (inst/examples/sir.cpp:126-198 through dust2/unfilter.hpp:188-229): the
 * gamma term multiplied by exp(-gamma) whatever dt (sir.cpp:175), no -adj_S
 * from S0 = N - I0 in the I0 gradient (sir.cpp:126-141), NaN for missing
 * data (sir.cpp:179-197).  exact = 1: the derivative of the computed
 * log-likelihood (no counterpart in the reference).  Invalidates the last
 * gradient. */
int dust2_unfilter_sir_set_adjoint_form(dust2_unfilter_sir* h, int exact, char* err,
                                        size_t errlen);

/* Replaces dust2_unfilter_update_pars<T> (dust2/r/unfilter.hpp:10-19): I0,
 * beta and gamma update, N and exp_noise are constant. */
int dust2_unfilter_sir_update_pars(dust2_unfilter_sir* h, const double* pars,
                                   const size_t* index_group, size_t n_index, char* err,
                                   size_t errlen);
/* out [G][P][S] */
int dust2_unfilter_sir_last_state(dust2_unfilter_sir* h, double* out, char* err,
                                  size_t errlen);
void dust2_unfilter_sir_free(dust2_unfilter_sir* h);

/* ---- serving: continuous batching of likelihood evaluations ---------------
 * No counterpart in the reference, whose MCMC driver evaluates one
 * likelihood at a time (R/monty.R:188-204).  Each request is one parameter
 * group's complete filter (sir, seir16: FP32) or ODE unfilter (sirode: FP64)
 * run on its own series.  Requests are admitted when their device footprint
 * fits the server's HBM budget (a first-fit arena over one device block,
 * uploads staged through one pinned host block: nothing is allocated per
 * request), stay resident, and advance chunk_intervals data intervals per
 * launch.  Requests of the same model and particle count share launches
 * (one HIP graph per group, replayed); each group runs on its own stream and
 * progresses at its own pace, its completion seen through an event.  A
 * request's result depends only on its own inputs and streams: it is
 * identical alone, in any mix and with the plain filter entry points on one
 * group. */
typedef struct dust2_server dust2_server;
#define DUST2_MODEL_SIR 0
#define DUST2_MODEL_SEIR16 1
#define DUST2_MODEL_SIRODE 2
int dust2_server_create(size_t hbm_bytes, int chunk_intervals, dust2_server** out, char* err,
                        size_t errlen);
/* pars: one parameter row of the model; data: n_time x n_data values.
 * chain < 0: the filter's streams are seeded from the integer `seed`
 * (stream k = jump^k of the splitmix64 expansion, as dust2_filter_sir_alloc).
 * chain >= 0: the request belongs to an MCMC chain whose streams stay
 * resident and continue from one request of the chain to the next, as a
 * chain's filter object does; the chain's first request seeds them at
 * long_jump^chain of the splitmix64 expansion of `seed`, then jumps
 * (R/monty.R:29-36).  A chain's next request waits while one is resident. */
int dust2_server_submit(dust2_server* s, int model, const double* pars, size_t n_particles,
                        double time_start, const double* time, size_t n_time, double dt,
                        const double* data, uint64_t seed, int64_t chain, uint64_t* id_out,
                        char* err, size_t errlen);
/* One scheduling pass: collects completed launches, admits queued requests,
 * launches idle groups; waits only when nothing could be done while work is
 * in flight.  Reports resident and queued requests. */
int dust2_server_step(dust2_server* s, size_t* n_active, size_t* n_queued, char* err,
                      size_t errlen);
/* Up to max finished results: id, log-likelihood, status (0 ok, 1 the
 * request reported errors; `err` then holds the first message). */
int dust2_server_poll(dust2_server* s, uint64_t* ids, double* ll, int* status, size_t max,
                      size_t* n_out, char* err, size_t errlen);
int dust2_server_memory(dust2_server* s, size_t* used, size_t* capacity, char* err,
                        size_t errlen);
void dust2_server_free(dust2_server* s);

/* Jump (2^128 draws) or, with long_jump != 0, long-jump (2^192 draws) each of
 * n_streams xoshiro256+ states in place (4 words each): monty's jump() and
 * long_jump() on exported streams (vignettes/design.Rmd:72-81,
 * R/monty.R:29-36). */
int dust2_amd_rng_jump(uint64_t* state, size_t n_streams, int long_jump, char* err,
                       size_t errlen);

/* ---- kernels exposed for parity tests --------------------------------------
 * Systematic resampling of one weight vector on the GPU with the filter's
 * own routines (replaces the reference test driver test_resample_weight /
 * test_scale_log_weights, src/test.cpp:10-23; dust2/filter_details.hpp:11-74):
 * one workgroup for n <= 2048, the multi-workgroup resampler of large
 * groups above (n <= 4096 * 8192).  log_weights: n log-weights; idx: n
 * ancestors; out3: [ll increment, resampled flag, sequential-fallback flag]. */
int dust2_amd_resample(const double* log_weights, size_t n, double u, uint32_t* idx,
                       double* out3, char* err, size_t errlen);

/* The same resampling with nothing allocated: no counterpart in the
 * reference; for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid allocation).
 * d_log_weights (n doubles), d_idx (n uint32) and d_out3 (3 doubles) are
 * the caller's device buffers; the work is queued on `stream` (a
 * hipStream_t, NULL for the default stream) and the call returns without
 * waiting.  The multi-workgroup path keeps its partial results in a
 * library-owned device array of 8.4 MB per GPU, so n <= 1,048,576 (larger n
 * is refused).  Calls that overlap on one GPU share that array and race:
 * order them on one stream or synchronise between them. */
int dust2_amd_resample_noalloc(const double* d_log_weights, size_t n, double u, uint32_t* d_idx,
                               double* d_out3, void* stream, char* err, size_t errlen);

/* Lineage export of given saved states (the kernel behind
 * dust2_filter_last_trajectories): raw [n_time][G][P][S], ancestry
 * [n_time][G][P] or NULL, select [G] or NULL; out as
 * dust2_filter_last_trajectories.  For the tests. */
int dust2_amd_trajectory_lineage(const double* raw, const uint32_t* ancestry, size_t n_time,
                                 size_t n_groups, size_t n_particles, size_t n_state,
                                 const size_t* select, double* out, char* err, size_t errlen);
/* Evaluates an interpolator of time-varying inputs on the device: kind
 * 0 piecewise constant, 1 piecewise linear, 2 natural cubic spline; times
 * t[n_t] strictly increasing, values y[n_t * n_series] (y[i * n_series + j]
 * for time i of series j); out[n_z * n_series] at the times z.  Replaces
 * dust2::interpolate::Interpolate{Constant,Linear,Spline}{,Array}::eval
 * (dust2/interpolate/interpolate.hpp:69-268, spline slopes
 * dust2/interpolate/spline.hpp:17-90), with the reference's errors for
 * times outside the knots.  For the tests and for model authors. */
int dust2_amd_interpolate(int kind, const double* t, size_t n_t, const double* y,
                          size_t n_y, size_t n_series, const double* z, size_t n_z, double* out,
                          char* err, size_t errlen);
/* n streams of raw xoshiro256+ state, the filter/system seeding path. */
int dust2_amd_rng_streams(const uint64_t* seed_state, size_t n_seed_streams, uint64_t seed,
                          size_t n_streams, uint64_t* out, char* err, size_t errlen);

/* Sampler draws on the GPU: kind 0 binomial(a=n, b=p), 1 exponential(a=rate),
 * 2 normal(a=mean, b=sd), 3 poisson(a=lambda), 4 random_real.  n_draws per
 * stream from each of n_streams streams (rng: [stream][4], updated);
 * out [stream][draw]; err_out per stream (0 or 1 = invalid binomial). */
int dust2_amd_sample(int kind, int precision, uint64_t* rng, size_t n_streams,
                     size_t n_draws, const double* a, const double* b, int deterministic,
                     double* out, uint32_t* err_out, char* err, size_t errlen);

/* Vector-ALU issue-rate calibration (no counterpart in the reference; it
 * supports the measurement of the sampler kernels, which are bound by
 * vector instructions, not HBM): waves_per_simd wavefronts per SIMD on
 * every compute unit, each issuing `iters` iterations of independent FP32
 * (precision DUST2_FP32: 8 per iteration) or FP64 (DUST2_FP64: 4) fused
 * multiply-adds.  ms_out: the kernel time (HIP events, after one warm-up
 * launch); rate_out: fused multiply-add wave-instructions per SIMD per
 * nanosecond; n_simd_out: the SIMDs of the device (4 per compute unit). */
int dust2_amd_valu_probe(int precision, int waves_per_simd, int iters, double* ms_out,
                         double* rate_out, int* n_simd_out, char* err, size_t errlen);

#ifdef __cplusplus
}
#endif

#endif
