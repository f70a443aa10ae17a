// Host runtime behind the C ABI (include/dust2_amd.h): argument checks with
// the reference's messages, device buffers owned by handles, RNG stream
// expansion, launch of the kernels and conversion between the caller's
// [group][particle][state] layout and the device's per-group SoA layout.
#include "../../include/dust2_amd.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <deque>
#include <map>
#include <tuple>
#include <cstdio>
#include <cstring>
#include <limits>
#include <memory>
#include <mutex>
#include <sstream>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "engine.hpp"
#include "interpolate.hpp"
#include "models.hpp"
#include "ode_general.hpp"
#include "ode_models.hpp"

using namespace dust2_amd;

namespace {

struct hip_error : std::runtime_error {
  using std::runtime_error::runtime_error;
};

void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) {
    throw hip_error(std::string("HIP error in ") + what + ": " + hipGetErrorString(e));
  }
}

int fail(const std::exception& e, char* err, size_t errlen) {
  if (err && errlen) std::snprintf(err, errlen, "%s", e.what());
  return 1;
}

// Device allocation owned by a handle.
class dbuf {
 public:
  dbuf() = default;
  explicit dbuf(size_t bytes) { alloc(bytes); }
  ~dbuf() { release(); }
  dbuf(const dbuf&) = delete;
  dbuf& operator=(const dbuf&) = delete;
  void alloc(size_t bytes) {
    release();
    if (bytes) hip_check(hipMalloc(&ptr_, bytes), "hipMalloc");
    bytes_ = bytes;
  }
  void release() {
    if (ptr_) (void)hipFree(ptr_);
    ptr_ = nullptr;
    bytes_ = 0;
  }
  template <typename T = void>
  T* get() const {
    return static_cast<T*>(ptr_);
  }
  size_t bytes() const { return bytes_; }
  void swap(dbuf& o) {
    std::swap(ptr_, o.ptr_);
    std::swap(bytes_, o.bytes_);
  }

 private:
  void* ptr_ = nullptr;
  size_t bytes_ = 0;
};

// Pinned host memory (hipHostMalloc), owned.
class pinned {
 public:
  pinned() = default;
  pinned(const pinned&) = delete;
  pinned& operator=(const pinned&) = delete;
  ~pinned() { release(); }
  void alloc(size_t bytes) {
    release();
    if (bytes) hip_check(hipHostMalloc(&p_, bytes, hipHostMallocDefault), "hipHostMalloc");
    n_ = bytes;
  }
  void release() {
    if (p_) (void)hipHostFree(p_);
    p_ = nullptr;
    n_ = 0;
  }
  template <typename T = void>
  T* get() const {
    return static_cast<T*>(p_);
  }
  size_t bytes() const { return n_; }

 private:
  void* p_ = nullptr;
  size_t n_ = 0;
};

void to_device(dbuf& d, const void* src, size_t bytes, hipStream_t s) {
  if (bytes) hip_check(hipMemcpyAsync(d.get(), src, bytes, hipMemcpyHostToDevice, s), "copy to device");
}

void to_host(void* dst, const dbuf& d, size_t bytes, hipStream_t s) {
  if (bytes) {
    hip_check(hipMemcpyAsync(dst, d.get(), bytes, hipMemcpyDeviceToHost, s), "copy to host");
  }
  hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
}

// ---- argument checks (dust2/r/helpers.hpp) --------------------------------
void check_dt(double dt) {
  // dt must be positive and an integer or 1/integer (helpers.hpp:232-262)
  if (!(dt > 0)) {
    throw std::runtime_error("'dt' must be greater than 0");
  }
  if (dt > 1) {
    if (std::fmod(dt, 1) != 0) {
      throw std::runtime_error("'dt' must be an integer or the inverse of an integer");
    }
  } else {
    const double inv = 1 / dt;
    if (std::fabs(inv - std::round(inv)) > 1e-8) {
      throw std::runtime_error("'dt' must be an integer or the inverse of an integer");
    }
  }
}

void check_time_sequence(double time_start, const double* time, size_t n) {
  if (n == 0) {
    throw std::runtime_error("Expected at least one value in 'time'");
  }
  if (!(time[0] > time_start)) {
    throw std::runtime_error("Expected 'time[1]' (" + std::to_string(time[0]) +
                             ") to be larger than the previous value (" +
                             std::to_string(time_start) + ")");
  }
  for (size_t i = 1; i < n; ++i) {
    if (!(time[i] > time[i - 1])) {
      throw std::runtime_error("Expected 'time' to be strictly increasing");
    }
  }
}

void check_sizes(size_t n_groups, size_t n_particles) {
  if (n_groups == 0) throw std::runtime_error("'n_groups' must be at least 1");
  if (n_particles == 0) throw std::runtime_error("'n_particles' must be at least 1");
  if (n_particles > (1u << 24) || n_groups > (1u << 24)) {
    throw std::runtime_error("'n_particles' or 'n_groups' is too large");
  }
}

std::vector<size_t> check_index(const size_t* index, size_t n_index, size_t n_groups) {
  std::vector<size_t> ret;
  if (index == nullptr) {
    for (size_t i = 0; i < n_groups; ++i) ret.push_back(i);
    return ret;
  }
  std::vector<char> seen(n_groups, 0);
  for (size_t i = 0; i < n_index; ++i) {
    if (index[i] >= n_groups) {
      throw std::runtime_error("Invalid index for 'index_group': values must lie in [0, " +
                               std::to_string(n_groups - 1) + "]");
    }
    // a repeated group would be run twice, concurrently, on one state;
    // dust_likelihood_run rejects duplicates (R/interface-likelihood.R:56,
    // R/filter-support.R:85-86)
    if (seen[index[i]]) {
      throw std::runtime_error("All elements of 'index_group' must be distinct");
    }
    seen[index[i]] = 1;
    ret.push_back(index[i]);
  }
  return ret;
}

// index_state of a run (0-based): the states its trajectories keep
// (trajectories<T>::set_index_and_reset, dust2/trajectories.hpp:31-45),
// checked as dust_likelihood_run does (R/interface-likelihood.R:97-98 with
// check_index, R/filter-support.R:73-91: in range, distinct).  Null keeps
// every state and gives an empty list.
std::vector<uint32_t> check_state_index(const size_t* index, size_t n_index, size_t S) {
  std::vector<uint32_t> ret;
  if (index == nullptr) return ret;
  if (n_index == 0) throw std::runtime_error("'index_state' must have nonzero length");
  std::vector<char> seen(S, 0);
  for (size_t i = 0; i < n_index; ++i) {
    if (index[i] >= S) {
      throw std::runtime_error("Invalid index for 'index_state': values must lie in [0, " +
                               std::to_string(S - 1) + "]");
    }
    if (seen[index[i]]) {
      throw std::runtime_error("All elements of 'index_state' must be distinct");
    }
    seen[index[i]] = 1;
    ret.push_back(static_cast<uint32_t>(index[i]));
  }
  return ret;
}

// save_snapshots of a run, checked as R/interface-likelihood.R:83-93 does
// (increasing, each one of the data times): per data time its snapshot
// slot or -1, and the number of snapshots in n_snap.
std::vector<int32_t> snapshot_slots(const double* snaps, size_t n, const std::vector<double>& times,
                                    size_t& n_snap) {
  if (snaps == nullptr) n = 0;
  for (size_t i = 1; i < n; ++i) {
    if (!(snaps[i] > snaps[i - 1])) {
      throw std::runtime_error("Values in 'save_snapshots' must be increasing");
    }
  }
  std::vector<int32_t> slot(times.size(), -1);
  for (size_t i = 0; i < n; ++i) {
    const auto it = std::find(times.begin(), times.end(), snaps[i]);
    if (it == times.end()) {
      throw std::runtime_error("All elements of 'save_snapshots' must be found in 'time'");
    }
    slot[static_cast<size_t>(it - times.begin())] = static_cast<int32_t>(i);
  }
  n_snap = n;
  return slot;
}

// What a run keeps beside its log-likelihoods (trajectories<T>,
// dust2/trajectories.hpp:31-57): trajectories of the states of
// index_state and the whole state at the snapshot times, both read from
// the states saved at every data time.  A run checks its request first,
// clears the record when it starts and stores the request when it
// completes, so a failed run leaves nothing current (dust2/filter.hpp:
// 95-99 marks trajectories current after the time loop).
struct saved_run {
  bool traj = false;             // trajectories current
  size_t n_snap = 0;             // snapshots current (0: none)
  std::vector<uint32_t> states;  // index_state (empty: every state)
  std::vector<int32_t> slot;     // per data time: snapshot slot or -1
  bool any() const { return traj || n_snap > 0; }
};

saved_run check_save(int save_trajectories, const double* save_snapshots, size_t n_snapshots,
                     const size_t* index_state, size_t n_index_state, size_t S,
                     const std::vector<double>& times) {
  saved_run r;
  r.traj = save_trajectories != 0;
  r.states = check_state_index(index_state, n_index_state, S);
  r.slot = snapshot_slots(save_snapshots, n_snapshots, times, r.n_snap);
  return r;
}

// The states an unfilter saved at every data time, device layout
// [t][S][n_total] (particle k of group g at k = g * P + p), as
// out[rows][groups][P][states]: rows are the data times t in order, or,
// with `slot`, the times with slot[t] >= 0 at row slot[t]; `states`
// empty means every state.
template <typename real, typename Group>
void export_saved_states(const real* saved, size_t T, size_t S, size_t n_total, size_t P,
                         const std::vector<Group>& groups, const std::vector<int32_t>* slot,
                         const std::vector<uint32_t>& states, double* out) {
  const size_t n_out = groups.size();
  const size_t S_out = states.empty() ? S : states.size();
  for (size_t t = 0; t < T; ++t) {
    const int64_t row = slot ? (*slot)[t] : static_cast<int64_t>(t);
    if (row < 0) continue;
    for (size_t i = 0; i < n_out; ++i) {
      for (size_t p = 0; p < P; ++p) {
        const size_t k = static_cast<size_t>(groups[i]) * P + p;
        double* o = out + ((static_cast<size_t>(row) * n_out + i) * P + p) * S_out;
        for (size_t c = 0; c < S_out; ++c) {
          const size_t src = states.empty() ? c : states[c];
          o[c] = static_cast<double>(saved[(t * S + src) * n_total + k]);
        }
      }
    }
  }
}

// ---- xoshiro256+ on the host (seeding only) --------------------------------
struct host_rng {
  uint64_t s[4];
};

uint64_t host_next(host_rng& r) {
  const uint64_t out = r.s[0] + r.s[3];
  const uint64_t t = r.s[1] << 17;
  r.s[2] ^= r.s[0];
  r.s[3] ^= r.s[1];
  r.s[1] ^= r.s[2];
  r.s[0] ^= r.s[3];
  r.s[2] ^= t;
  r.s[3] = (r.s[3] << 45) | (r.s[3] >> 19);
  return out;
}

// Published xoshiro256 jump polynomial (2^128 steps).
constexpr uint64_t kJumpPoly[4] = {0x180ec6d33cfd0abaULL, 0xd5a61266f0c9392cULL,
                                   0xa9582618e03fc9aaULL, 0x39abdc4529b1661cULL};

// The long-jump polynomial of xoshiro256 (2^192 draws; Blackman & Vigna's
// published constants): the per-chain start of monty's seeding
// (R/monty.R:29-36, vignettes/design.Rmd:81).
constexpr uint64_t kLongJumpPoly[4] = {0x76e15d3efefdcbbfULL, 0xc5004e441c522fb3ULL,
                                       0x77710069854ee241ULL, 0x39109bb02acbe635ULL};

void host_apply_poly(host_rng& r, const uint64_t* poly) {
  uint64_t acc[4] = {0, 0, 0, 0};
  for (int w = 0; w < 4; ++w) {
    for (int b = 0; b < 64; ++b) {
      if (poly[w] & (uint64_t(1) << b)) {
        for (int i = 0; i < 4; ++i) acc[i] ^= r.s[i];
      }
      host_next(r);
    }
  }
  std::memcpy(r.s, acc, sizeof(acc));
}

void host_jump(host_rng& r) { host_apply_poly(r, kJumpPoly); }
void host_long_jump(host_rng& r) { host_apply_poly(r, kLongJumpPoly); }

// splitmix64-based expansion of an integer seed into the first stream.
void host_seed_words(uint64_t seed, uint64_t* out) {
  for (int i = 0; i < 4; ++i) {
    uint64_t z = seed + 0x9e3779b97f4a7c15ULL;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    seed = z ^ (z >> 31);
    out[i] = seed;
  }
}

constexpr int kJumpMats = 32;  // supports 2^32 streams

// J_i = jump^(2^i) as column-major GF(2) matrices: column j = J_i e_j.
const std::vector<uint64_t>& host_jump_matrices() {
  static std::vector<uint64_t> mats;
  static std::once_flag once;
  std::call_once(once, [] {
    mats.assign(static_cast<size_t>(kJumpMats) * 256 * 4, 0);
    for (int j = 0; j < 256; ++j) {
      host_rng r{{0, 0, 0, 0}};
      r.s[j / 64] = uint64_t(1) << (j % 64);
      host_jump(r);
      std::memcpy(&mats[j * 4], r.s, 32);
    }
    for (int i = 1; i < kJumpMats; ++i) {
      const uint64_t* prev = &mats[static_cast<size_t>(i - 1) * 1024];
      uint64_t* cur = &mats[static_cast<size_t>(i) * 1024];
      for (int j = 0; j < 256; ++j) {
        const uint64_t* v = prev + j * 4;  // J_{i-1} e_j
        uint64_t acc[4] = {0, 0, 0, 0};
        for (int b = 0; b < 256; ++b) {
          if ((v[b / 64] >> (b % 64)) & 1u) {
            for (int w = 0; w < 4; ++w) acc[w] ^= prev[b * 4 + w];
          }
        }
        std::memcpy(cur + j * 4, acc, 32);
      }
    }
  });
  return mats;
}

struct device_tables {
  dbuf jump;
};

device_tables& tables_for_current_device() {
  static std::mutex mu;
  static std::vector<std::unique_ptr<device_tables>> per_device;
  int dev = 0;
  hip_check(hipGetDevice(&dev), "hipGetDevice");
  std::lock_guard<std::mutex> lock(mu);
  if (per_device.size() <= static_cast<size_t>(dev)) per_device.resize(dev + 1);
  if (!per_device[dev]) {
    auto t = std::make_unique<device_tables>();
    const auto& m = host_jump_matrices();
    t->jump.alloc(m.size() * sizeof(uint64_t));
    hip_check(hipMemcpy(t->jump.get(), m.data(), m.size() * sizeof(uint64_t),
                        hipMemcpyHostToDevice),
              "upload jump tables");
    per_device[dev] = std::move(t);
  }
  return *per_device[dev];
}

// Fill d_aos (n_streams x 4 words) with the seeded streams.
void expand_streams(const uint64_t* seed_state, size_t n_seed, uint64_t seed, size_t n_streams,
                    uint64_t* d_aos, hipStream_t s) {
  if (n_streams > (uint64_t(1) << kJumpMats)) {
    throw std::runtime_error("too many random number streams");
  }
  size_t n_copy = 0;
  uint64_t s0[4];
  if (seed_state && n_seed > 0) {
    n_copy = std::min(n_seed, n_streams);
    hip_check(hipMemcpyAsync(d_aos, seed_state, n_copy * 32, hipMemcpyHostToDevice, s),
              "copy seed");
    host_rng r;
    std::memcpy(r.s, seed_state + 4 * (n_seed - 1), 32);
    host_jump(r);
    std::memcpy(s0, r.s, 32);
  } else {
    host_seed_words(seed, s0);
  }
  if (n_copy < n_streams) {
    hip_check(launch_seed_streams(s0, tables_for_current_device().jump.get<uint64_t>(),
                                  kJumpMats, n_streams - n_copy, d_aos + 4 * n_copy, 0, s),
              "seed streams");
  }
}

// ---- error reporting (dust2/errors.hpp:34-55) --------------------------------
std::string error_text(const error_record& r) {
  char buf[256];
  switch (r.code) {
    case ERR_BINOMIAL:
      std::snprintf(buf, sizeof(buf), "Invalid call to binomial with n = %.0f, p = %g, q = %g",
                    r.a, r.b, 1 - r.b);
      return buf;
    case ERR_ODE_MAX_STEPS: return "too many steps";
    case ERR_ODE_STEP_SMALL: return "step too small";
    case ERR_ODE_VANISHED: return "step size vanished";
    case ERR_ODE_INIT: return "Initial step size was not finite";
    case ERR_ODE_HISTORY_LOCATE: return "Failed to locate history element";
    case ERR_ODE_HISTORY_SHORT: {
      std::ostringstream msg;
      msg << "History too short for a delayed value at time " << r.a
          << ": raise 'history_cap' (steps kept per particle)";
      return msg.str();
    }
    case ERR_INTERP_BEFORE:
    case ERR_INTERP_AFTER: {
      // dust2/interpolate/interpolate.hpp:20-40; a = time, b = first or last knot
      std::ostringstream msg;
      msg << "Tried to interpolate at time = " << r.a << ", which is "
          << (r.code == ERR_INTERP_BEFORE ? r.b - r.a : r.a - r.b)
          << (r.code == ERR_INTERP_BEFORE ? " before the first time (" : " after the last time (")
          << r.b << ")";
      return msg.str();
    }
    default: return "unknown error";
  }
}

// records: one per particle, in particle order; count = number set.
std::string error_report(const std::vector<error_record>& records, size_t count,
                         const char* title = "particles", size_t n_max = 4) {
  std::stringstream msg;
  msg << count << " " << title << " reported errors.";
  const size_t n_report = std::min(n_max, count);
  for (size_t i = 0, j = 0; i < records.size() && j < n_report; ++i) {
    if (records[i].code != ERR_NONE) {
      msg << std::endl << "  - " << i + 1 << ": " << error_text(records[i]);
      ++j;
    }
  }
  if (n_report < count) {
    msg << std::endl << "  - (and " << (count - n_report) << " more)";
  }
  return msg.str();
}

int max_lds_per_block() {
  int dev = 0;
  int v = 0;
  hip_check(hipGetDevice(&dev), "hipGetDevice");
  hip_check(hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev),
            "hipDeviceGetAttribute");
  return v;
}

// SIR parameter row -> kernel parameter block (pars[8]):
// N, I0, beta, gamma, exp_noise, p_IR (fp64), p_IR (fp32).
void sir_kernel_pars(const double* row, double dt, double* out) {
  for (int i = 0; i < 5; ++i) out[i] = row[i];
  // FP64: the reference's expression (inst/examples/sir.cpp:64); FP32: the
  // same probability without the cancellation of 1 - exp(-x) in float
  out[5] = 1 - std::exp(-row[3] * dt);
  out[6] = static_cast<double>(static_cast<float>(-std::expm1(-row[3] * dt)));
  out[7] = 0;
}

constexpr size_t kSirPars = 5;
constexpr int kSirState = 5;

void check_sir_pars(const double* row) {
  const char* names[] = {"N", "I0", "beta", "gamma", "exp_noise"};
  for (size_t i = 0; i < kSirPars; ++i) {
    if (!std::isfinite(row[i])) {
      throw std::runtime_error(std::string("Parameter '") + names[i] + "' must be finite");
    }
  }
}


void check_finite_pars(const double* row, size_t n) {
  for (size_t i = 0; i < n; ++i) {
    if (!std::isfinite(row[i])) {
      throw std::runtime_error("Parameter " + std::to_string(i) + " must be finite");
    }
  }
}

// SEIR16 host parameter row: beta, sigma, gamma, exp_noise, I0, N[16],
// C[16][16] (row-major).  Kernel block: seir16_model in models.hpp.
constexpr size_t kSeirAge = 16;
constexpr size_t kSeirPars = 5 + kSeirAge + kSeirAge * kSeirAge;
constexpr size_t kSeirKpars = seir16_model::kpars;

void seir_kernel_pars(const double* row, double dt, double* out) {
  std::fill(out, out + kSeirKpars, 0.0);
  const double beta = row[0];
  out[0] = beta;
  out[1] = static_cast<float>(-std::expm1(-row[1] * dt));
  out[2] = static_cast<float>(-std::expm1(-row[2] * dt));
  out[3] = row[3];
  out[4] = row[4];
  const double* N = row + 5;
  const double* C = row + 5 + kSeirAge;
  for (size_t a = 0; a < kSeirAge; ++a) {
    out[5 + a] = N[a];
    for (size_t b = 0; b < kSeirAge; ++b) {
      out[seir16_model::kParA + a * kSeirAge + b] = beta * C[a * kSeirAge + b] / N[b];
      out[seir16_model::kParC + a * kSeirAge + b] = C[a * kSeirAge + b];
    }
  }
  // FP64: the reference-style expression 1 - exp(-rate * dt)
  out[seir16_model::kParP64] = 1 - std::exp(-row[1] * dt);
  out[seir16_model::kParP64 + 1] = 1 - std::exp(-row[2] * dt);
}


// Random walk (inst/examples/walk.cpp): host row sd, len, random_initial;
// the kernel block is the row.
constexpr size_t kWalkPars = 3;

void check_walk_pars(const double* row) {
  if (!std::isfinite(row[0])) throw std::runtime_error("Parameter 'sd' must be finite");
  if (!(row[1] >= 1 && row[1] <= walk_model::max_len && row[1] == std::floor(row[1]))) {
    throw std::runtime_error("Parameter 'len' must be an integer in [1, " +
                             std::to_string(walk_model::max_len) + "]");
  }
  if (row[2] != 0 && row[2] != 1) {
    throw std::runtime_error("Parameter 'random_initial' must be TRUE or FALSE");
  }
}

void walk_kernel_pars(const double* row, double, double* out) {
  std::copy(row, row + kWalkPars, out);
}

// Host side of each discrete-time model: its parameter row, validation,
// kernel parameter block (models.hpp) and sizes.  The kernels themselves
// are one set templated on the model.
struct discrete_model_host {
  int id;
  const char* name;
  size_t n_pars;     // host parameter row
  size_t n_data;     // data values per time (0: the model has no compare_data)
  size_t kpars;      // kernel parameter block (doubles)
  double zero_period;
  size_t (*n_state)(const double* row);
  void (*check)(const double* row);
  void (*kernel_pars)(const double* row, double dt, double* out);
  // the parameters update_pars may change (others are constant,
  // dust2/properties.hpp update_shared)
  std::vector<size_t> updatable;
};

const discrete_model_host& discrete_model(int id) {
  static const discrete_model_host sir{
      MODEL_SIR, "sir", kSirPars, 1, 8, 1.0, [](const double*) -> size_t { return 5; },
      check_sir_pars, sir_kernel_pars,
      // I0, beta, gamma; N and exp_noise are constant (inst/examples/sir.cpp:89-93)
      {1, 2, 3}};
  static const discrete_model_host seir{
      MODEL_SEIR16, "seir16", kSeirPars, kSeirAge, kSeirKpars, 7.0,
      [](const double*) -> size_t { return 5 * kSeirAge; },
      [](const double* row) { check_finite_pars(row, kSeirPars); }, seir_kernel_pars,
      // beta, sigma, gamma, I0; exp_noise, N and C are constant
      {0, 1, 2, 4}};
  static const discrete_model_host walk{
      MODEL_WALK, "walk", kWalkPars, 0, kWalkPars, 1.0,
      [](const double* row) -> size_t { return static_cast<size_t>(row[1]); },
      check_walk_pars, walk_kernel_pars,
      // sd (walk.cpp update_shared); len and random_initial are constant
      {0}};
  switch (id) {
    case MODEL_SIR: return sir;
    case MODEL_SEIR16: return seir;
    case MODEL_WALK: return walk;
    default: throw std::runtime_error("unknown discrete-time model");
  }
}

int discrete_model_by_name(const char* name) {
  for (int id : {MODEL_SIR, MODEL_SEIR16, MODEL_WALK}) {
    if (name && std::strcmp(name, discrete_model(id).name) == 0) return id;
  }
  throw std::runtime_error(std::string("unknown discrete-time model '") + (name ? name : "") +
                           "'");
}

// Every group must have the state length of the first
// (dust2/r/helpers.hpp:333-369).
size_t discrete_state_size(const discrete_model_host& m, const double* pars, size_t n_groups) {
  const size_t S = m.n_state(pars);
  for (size_t g = 1; g < n_groups; ++g) {
    const size_t Sg = m.n_state(pars + g * m.n_pars);
    if (Sg != S) {
      throw std::runtime_error("State length for group " + std::to_string(g + 1) +
                               " was different to previous groups; total length was expected "
                               "to be " + std::to_string(S) + " but it was " +
                               std::to_string(Sg));
    }
  }
  return S;
}
}  // namespace

// ============================================================================
// Filter
// ============================================================================
// Every handle begins with a tag naming its type, so that an entry point
// handed another type of handle (through an untyped foreign-function
// pointer) reports it instead of reading the wrong object.
enum handle_tag : uint32_t {
  TAG_FILTER = 0x64326631u,
  TAG_SYSTEM = 0x64327331u,
  TAG_UNFILTER = 0x64327531u,
  TAG_UNFILTER_SIR = 0x64327532u,
  TAG_ODE_SYSTEM = 0x64327332u,
  TAG_ODE_GENERAL = 0x64327333u,
  TAG_FILTER_ODE = 0x64326632u,
  TAG_UNFILTER_ODE = 0x64327534u,
  TAG_SERVER = 0x64327631u,
};

template <typename H>
bool handle_ok(const H* h, uint32_t tag) {
  return h != nullptr && h->tag == tag;
}

template <typename H>
void check_handle(const H* h, uint32_t tag, const char* what) {
  if (h == nullptr) {
    throw std::runtime_error(std::string("Pointer has been serialised, cannot continue safely (") +
                             what + ")");
  }
  if (h->tag != tag) throw std::runtime_error(std::string("Not a ") + what + " handle");
}

struct dust2_filter {
  uint32_t tag = TAG_FILTER;
  int model = MODEL_SIR;     // discrete_model_id of models.hpp
  size_t S = 5;              // states per particle
  size_t n_pars = 5;         // host parameter row length
  size_t n_data = 1;         // data values per group per time
  size_t kpars = 8;          // kernel parameter block (doubles)
  double zero_period = 1;    // zero_every period of the model
  bool fp64 = true;
  size_t G = 0, P = 0, n_time = 0, n_groups_data = 0;
  double time_start = 0, dt = 1;
  std::vector<double> pars;  // G x 5 (host rows)
  std::vector<double> times;
  hipStream_t stream = nullptr;
  launch_config cfg{0, 0};
  int engine = DUST2_ENGINE_AUTO;
  bool use_graph = true;
  bool profile = false;
  dbuf d_times, d_obs, d_lfact, d_state, d_rng, d_ll, d_time, d_slots, d_errors, d_counters,
      d_ancestry, d_traj;
  // interval pipeline
  dbuf d_state_b, d_lambda, d_anc, d_sched_time, d_sched_steps, d_sched_zero, d_kpars, d_pslots,
      d_pslots_init;
  // groups above one workgroup's LDS resample across workgroups
  // (resample_kernels.hip): one task per slot, partial results in scratch
  bool multi_block = false;
  dbuf d_mb_tasks, d_mb_scratch;
  hipGraphExec_t graph = nullptr;
  std::vector<uint64_t> graph_key;
  // the pipeline splits the groups into `pipe_chains` (at most kMaxChains)
  // independent launch chains on their own streams, so one chain's
  // resample overlaps another chain's advance; chain c > 0 runs on
  // chain_stream[c - 1] and counts its errors in counter 1 + c
  static constexpr int kMaxChains = 4;
  int pipe_chains = 2;
  // threads per resample workgroup (1024, 512 or 256; the particles per
  // thread follow from the group size)
  int pipe_resample_block = 1024;
  hipStream_t chain_stream[kMaxChains - 1] = {};
  hipEvent_t ev_fork = nullptr;
  hipEvent_t ev_join[kMaxChains - 1] = {};
  bool traj_valid = false;
  bool ancestry_valid = false;
  std::vector<size_t> last_index;        // groups of the last run
  std::vector<uint32_t> traj_states;     // index_state of the last run (empty: all)
  std::vector<int32_t> snap_slot;        // per data time: snapshot slot or -1
  size_t n_snap = 0;                     // snapshots of the last run
  std::vector<uint8_t> select_valid;     // per group: a random particle is chosen
  // beta interpolated in time (SIR, pipeline engine): kind, knots shared by
  // the groups, values and spline slopes per group (or shared)
  int beta_kind = -1;
  size_t beta_n = 0, beta_groups = 0;
  dbuf d_beta_t, d_beta_y, d_beta_k;
  dbuf d_select, d_groups, d_traj_out, d_traj_states;   // lineage export
  double last_ms = 0;
  double last_fallbacks = 0;
  double last_advance_ms = 0;   // mean advance-kernel time (profile mode)
  double last_resample_ms = 0;  // mean resample-kernel time (profile mode)
  int last_engine = 0;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;

  size_t real_size() const { return fp64 ? 8 : 4; }

  ~dust2_filter() {
    if (graph) (void)hipGraphExecDestroy(graph);
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (ev_fork) (void)hipEventDestroy(ev_fork);
    for (int c = 0; c < kMaxChains - 1; ++c) {
      if (ev_join[c]) (void)hipEventDestroy(ev_join[c]);
      if (chain_stream[c]) (void)hipStreamDestroy(chain_stream[c]);
    }
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace {

hipError_t launch_advance(const dust2_filter* h, const pipe_slot* slots, int n_slots, int P,
                          int t, int parity, uint32_t mode, hipStream_t s) {
  return launch_discrete_advance(h->model, h->fp64, slots, n_slots, P, t, parity, mode, s);
}


void aos_to_soa(const double* src, void* dst, size_t G, size_t P, size_t S, bool fp64) {
  for (size_t g = 0; g < G; ++g) {
    for (size_t p = 0; p < P; ++p) {
      for (size_t c = 0; c < S; ++c) {
        const double v = src[(g * P + p) * S + c];
        const size_t o = (g * S + c) * P + p;
        if (fp64) {
          static_cast<double*>(dst)[o] = v;
        } else {
          static_cast<float*>(dst)[o] = static_cast<float>(v);
        }
      }
    }
  }
}

void soa_to_aos(const void* src, double* dst, size_t G, size_t P, size_t S, bool fp64) {
  for (size_t g = 0; g < G; ++g) {
    for (size_t p = 0; p < P; ++p) {
      for (size_t c = 0; c < S; ++c) {
        const size_t o = (g * S + c) * P + p;
        dst[(g * P + p) * S + c] = fp64 ? static_cast<const double*>(src)[o]
                                        : static_cast<const float*>(src)[o];
      }
    }
  }
}

void filter_upload_data(dust2_filter* h, const double* data) {
  const size_t n = h->n_time * h->n_groups_data * h->n_data;
  std::vector<double> lfact(n);
  for (size_t i = 0; i < n; ++i) {
    lfact[i] = std::isnan(data[i]) ? 0.0 : std::lgamma(data[i] + 1);
  }
  h->d_obs.alloc(n * sizeof(double));
  h->d_lfact.alloc(n * sizeof(double));
  to_device(h->d_obs, data, n * sizeof(double), h->stream);
  to_device(h->d_lfact, lfact.data(), n * sizeof(double), h->stream);
}

void check_filter(dust2_filter* h) { check_handle(h, TAG_FILTER, "filter"); }

// Step schedule of the data intervals, as dust_discrete::run_to_time
// computes it (dust2/discrete/system.hpp:61-92): n_steps =
// round(max(0, t - time)/dt), time += n_steps*dt, and the zero_every rule
// fmod(time + s*dt, 1) == 0 evaluated once per step on the host.
void filter_setup_pipeline(dust2_filter* h) {
  if (h->d_sched_time.bytes()) return;
  const size_t T = h->n_time, G = h->G, P = h->P;
  std::vector<double> st(T);
  std::vector<uint32_t> ns(T);
  std::vector<uint64_t> zm(T);
  double tcur = h->time_start;
  for (size_t t = 0; t < T; ++t) {
    const size_t n = static_cast<size_t>(std::round(std::max(0.0, h->times[t] - tcur) / h->dt));
    st[t] = tcur;
    ns[t] = static_cast<uint32_t>(n);
    uint64_t m = 0;
    for (size_t s = 0; s < std::min<size_t>(n, 64); ++s) {
      if (std::fmod(tcur + s * h->dt, h->zero_period) == 0) m |= uint64_t(1) << s;
    }
    zm[t] = m;
    tcur = tcur + n * h->dt;
  }
  h->d_sched_time.alloc(T * sizeof(double));
  h->d_sched_steps.alloc(T * sizeof(uint32_t));
  h->d_sched_zero.alloc(T * sizeof(uint64_t));
  to_device(h->d_sched_time, st.data(), T * sizeof(double), h->stream);
  to_device(h->d_sched_steps, ns.data(), T * sizeof(uint32_t), h->stream);
  to_device(h->d_sched_zero, zm.data(), T * sizeof(uint64_t), h->stream);
  h->d_state_b.alloc(G * h->S * P * h->real_size());
  h->d_lambda.alloc(G * P * sizeof(double));
  h->d_anc.alloc(G * P * sizeof(uint32_t));
  h->d_kpars.alloc(G * h->kpars * sizeof(double));
  h->d_pslots.alloc(G * sizeof(pipe_slot));
  h->d_pslots_init.alloc(G * sizeof(pipe_slot));
  hip_check(hipStreamSynchronize(h->stream), "sync");
}

pipe_slot make_pipe_slot(dust2_filter* h, size_t g, bool save_traj) {
  const size_t P = h->P;
  const size_t rsz = h->real_size();
  pipe_slot sl{};
  sl.pars = h->d_kpars.get<double>() + h->kpars * g;
  const size_t goff = (h->n_groups_data == 1 ? 0 : g) * h->n_data;
  sl.obs = h->d_obs.get<double>() + goff;
  sl.lfact = h->d_lfact.get<double>() + goff;
  sl.obs_stride = static_cast<uint32_t>(h->n_groups_data * h->n_data);
  sl.n_particles = static_cast<uint32_t>(P);
  sl.flags = 0u;
  sl.group = static_cast<uint32_t>(g);
  sl.sched_time = h->d_sched_time.get<double>();
  sl.sched_steps = h->d_sched_steps.get<uint32_t>();
  sl.sched_zero = h->d_sched_zero.get<uint64_t>();
  sl.dt = h->dt;
  sl.state_a = h->d_state.get<char>() + g * h->S * P * rsz;
  sl.state_b = h->d_state_b.get<char>() + g * h->S * P * rsz;
  sl.rng = h->d_rng.get<uint64_t>() + g * 4 * (P + 2);
  sl.lambda = h->d_lambda.get<double>() + g * P;
  sl.ancestors = h->d_anc.get<uint32_t>() + g * P;
  sl.ll = h->d_ll.get<double>() + g;
  sl.ancestry = save_traj ? h->d_ancestry.get<uint32_t>() + g * h->n_time * P : nullptr;
  sl.traj = save_traj ? h->d_traj.get<char>() + g * h->n_time * h->S * P * rsz : nullptr;
  sl.errors = h->d_errors.get<error_record>() + g * P;
  sl.error_count = h->d_counters.get<uint32_t>();
  sl.n_fallback = h->d_counters.get<uint32_t>() + 1;
  if (h->beta_n > 0) {
    const size_t off = h->beta_groups == 1 ? 0 : g * h->beta_n;
    sl.beta_t = h->d_beta_t.get<double>();
    sl.beta_y = h->d_beta_y.get<double>() + off;
    sl.beta_k = h->d_beta_k.get<double>() + off;
    sl.beta_n = static_cast<int32_t>(h->beta_n);
    sl.beta_kind = h->beta_kind;
  }
  return sl;
}

// The launches of one pipeline run.  With `profile`, events around every
// launch accumulate per-kernel times (not usable under graph capture).
// Number of launch chains for n_slots groups: chain c runs slots
// [c * n / chains, (c + 1) * n / chains).  Profiling uses one chain so that
// every launch can be timed.
int pipeline_chains(const dust2_filter* h, int n_slots) {
  if (h->profile || n_slots < 2) return 1;
  return std::max(1, std::min({h->pipe_chains, dust2_filter::kMaxChains, n_slots}));
}

void chain_range(int n_slots, int chains, int c, int& first, int& count) {
  first = static_cast<int>(static_cast<long>(n_slots) * c / chains);
  count = static_cast<int>(static_cast<long>(n_slots) * (c + 1) / chains) - first;
}

// The launches of one pipeline run.  With `profile`, events around every
// launch accumulate per-kernel times (not usable under graph capture).
void pipeline_launches(dust2_filter* h, int n_slots, bool initial_given, bool save_traj,
                       std::vector<hipEvent_t>* evs) {
  const int P = static_cast<int>(h->P);
  const int T = static_cast<int>(h->n_time);
  const int chains = pipeline_chains(h, n_slots);
  if (chains > 1) {
    // fork: the other chains start after everything queued before them
    hip_check(hipEventRecord(h->ev_fork, h->stream), "hipEventRecord");
    for (int c = 1; c < chains; ++c) {
      hip_check(hipStreamWaitEvent(h->chain_stream[c - 1], h->ev_fork, 0), "hipStreamWaitEvent");
    }
  }
  for (int c = 0; c < chains; ++c) {
    hipStream_t s = c == 0 ? h->stream : h->chain_stream[c - 1];
    int first = 0, count = 0;
    chain_range(n_slots, chains, c, first, count);
    const pipe_slot* slots = h->d_pslots.get<pipe_slot>() + first;
    auto mark = [&]() {
      if (evs) {
        hipEvent_t e;
        hip_check(hipEventCreate(&e), "hipEventCreate");
        hip_check(hipEventRecord(e, s), "hipEventRecord");
        evs->push_back(e);
      }
    };
    hip_check(launch_pipe_reset(slots, count, s), "reset kernel");
    int parity = 0;
    for (int t = 0; t < T; ++t) {
      uint32_t mode = PIPE_STEP;
      if (t == 0) {
        mode |= initial_given ? 0u : PIPE_INITIAL;
      } else {
        mode |= PIPE_GATHER | (save_traj ? PIPE_TRAJ_PREV : 0u);
      }
      mark();
      hip_check(launch_advance(h, slots, count, P, t, parity, mode, s), "advance kernel");
      mark();
      parity ^= 1;
      if (h->multi_block) {
        hip_check(launch_mb_resample(h->d_mb_tasks.get<mb_resample>() + first, count, P, t, s),
                  "resample kernels");
      } else {
        hip_check(launch_pipe_resample(slots, count, P, t, s, h->pipe_resample_block),
                  "resample kernel");
      }
      mark();
    }
    hip_check(launch_advance(h, slots, count, P, T, parity,
                             PIPE_GATHER | (save_traj ? PIPE_TRAJ_PREV : 0u), s),
              "gather kernel");
  }
  for (int c = 1; c < chains; ++c) {
    // join
    hip_check(hipEventRecord(h->ev_join[c - 1], h->chain_stream[c - 1]), "hipEventRecord");
    hip_check(hipStreamWaitEvent(h->stream, h->ev_join[c - 1], 0), "hipStreamWaitEvent");
  }
}

// Runs the filter through the interval pipeline; the final state is left
// in d_state.
void pipeline_run(dust2_filter* h, const std::vector<size_t>& idx, bool initial_given,
                  bool save_traj) {
  filter_setup_pipeline(h);
  hipStream_t s = h->stream;
  const size_t G = h->G;
  std::vector<double> kp(G * h->kpars);
  const discrete_model_host& m = discrete_model(h->model);
  for (size_t g = 0; g < G; ++g) m.kernel_pars(&h->pars[g * h->n_pars], h->dt, &kp[g * h->kpars]);
  to_device(h->d_kpars, kp.data(), kp.size() * sizeof(double), s);
  std::vector<pipe_slot> slots;
  std::vector<char> selected(G, 0);
  for (size_t g : idx) {
    slots.push_back(make_pipe_slot(h, g, save_traj));
    selected[g] = 1;
  }
  {
    // each chain counts its own errors (counter 0, or 1 + c for chain c)
    // so that a failing chain stops only itself; the report keeps the
    // earliest interval
    const int chains = pipeline_chains(h, static_cast<int>(slots.size()));
    for (int c = 1; c < chains; ++c) {
      int first = 0, count = 0;
      chain_range(static_cast<int>(slots.size()), chains, c, first, count);
      for (int i = first; i < first + count; ++i) {
        slots[i].error_count = h->d_counters.get<uint32_t>() + 1 + c;
      }
    }
  }
  to_device(h->d_pslots, slots.data(), slots.size() * sizeof(pipe_slot), s);
  if (h->multi_block) {
    const size_t per = mb_resample_scratch_doubles(static_cast<int>(h->P));
    if (h->d_mb_scratch.bytes() < G * per * sizeof(double)) {
      h->d_mb_scratch.alloc(G * per * sizeof(double));
    }
    if (h->d_mb_tasks.bytes() < G * sizeof(mb_resample)) {
      h->d_mb_tasks.alloc(G * sizeof(mb_resample));
    }
    std::vector<mb_resample> tasks(slots.size());
    for (size_t i = 0; i < slots.size(); ++i) {
      const pipe_slot& sl = slots[i];
      mb_resample& tk = tasks[i];
      tk.w = sl.lambda;
      tk.idx = sl.ancestors;
      tk.ancestry = sl.ancestry;
      tk.ll = sl.ll;
      tk.frng = sl.rng;
      tk.frng_stride = static_cast<uint32_t>(h->P + 2);
      tk.n_fallback = sl.n_fallback;
      tk.scratch = h->d_mb_scratch.get<double>() + sl.group * per;
      tk.P = static_cast<int>(h->P);
    }
    to_device(h->d_mb_tasks, tasks.data(), tasks.size() * sizeof(mb_resample), s);
  }
  hip_check(hipStreamSynchronize(s), "sync");
  const int n_slots = static_cast<int>(slots.size());
  const int T = static_cast<int>(h->n_time);

  if (h->profile) {
    std::vector<hipEvent_t> evs;
    hip_check(hipEventRecord(h->ev0, s), "hipEventRecord");
    pipeline_launches(h, n_slots, initial_given, save_traj, &evs);
    hip_check(hipEventRecord(h->ev1, s), "hipEventRecord");
    hip_check(hipStreamSynchronize(s), "sync");
    double adv = 0, res = 0;
    for (int t = 0; t < T; ++t) {
      float a = 0, b = 0;
      hip_check(hipEventElapsedTime(&a, evs[3 * t], evs[3 * t + 1]), "elapsed");
      hip_check(hipEventElapsedTime(&b, evs[3 * t + 1], evs[3 * t + 2]), "elapsed");
      adv += a;
      res += b;
    }
    for (auto e : evs) (void)hipEventDestroy(e);
    h->last_advance_ms = adv / T;
    h->last_resample_ms = res / T;
  } else if (h->use_graph) {
    const std::vector<uint64_t> key = {static_cast<uint64_t>(n_slots), h->P, h->n_time,
                                       initial_given, save_traj, h->fp64,
                                       static_cast<uint64_t>(pipeline_chains(h, n_slots)),
                                       static_cast<uint64_t>(h->pipe_resample_block)};
    if (!h->graph || key != h->graph_key) {
      if (h->graph) {
        hip_check(hipGraphExecDestroy(h->graph), "hipGraphExecDestroy");
        h->graph = nullptr;
      }
      hipGraph_t g;
      hip_check(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "begin capture");
      pipeline_launches(h, n_slots, initial_given, save_traj, nullptr);
      hip_check(hipStreamEndCapture(s, &g), "end capture");
      hip_check(hipGraphInstantiate(&h->graph, g, nullptr, nullptr, 0), "graph instantiate");
      hip_check(hipGraphDestroy(g), "hipGraphDestroy");
      h->graph_key = key;
    }
    hip_check(hipEventRecord(h->ev0, s), "hipEventRecord");
    hip_check(hipGraphLaunch(h->graph, s), "hipGraphLaunch");
    hip_check(hipEventRecord(h->ev1, s), "hipEventRecord");
  } else {
    hip_check(hipEventRecord(h->ev0, s), "hipEventRecord");
    pipeline_launches(h, n_slots, initial_given, save_traj, nullptr);
    hip_check(hipEventRecord(h->ev1, s), "hipEventRecord");
  }
  // the final gather wrote buffer (T + 1) % 2 in the A/B naming
  if ((T + 1) % 2 == 1) {
    if (initial_given) {
      // groups outside index_group keep the caller's initial state, which
      // is in buffer A; carry it over to B before B becomes the state
      const size_t bytes = h->S * h->P * h->real_size();
      for (size_t g = 0; g < G; ++g) {
        if (selected[g]) continue;
        hip_check(hipMemcpyAsync(h->d_state_b.get<char>() + g * bytes,
                                 h->d_state.get<char>() + g * bytes, bytes,
                                 hipMemcpyDeviceToDevice, s),
                  "hipMemcpyAsync");
      }
    }
    hip_check(hipStreamSynchronize(s), "sync");
    h->d_state.swap(h->d_state_b);
  }
  // groups outside index_group are reset to the initial state
  // (dust2/filter.hpp:248-250) but not run
  if (!initial_given) {
    std::vector<pipe_slot> rest;
    for (size_t g = 0; g < G; ++g) {
      if (!selected[g]) rest.push_back(make_pipe_slot(h, g, false));
    }
    if (!rest.empty()) {
      to_device(h->d_pslots_init, rest.data(), rest.size() * sizeof(pipe_slot), s);
      hip_check(launch_advance(h, h->d_pslots_init.get<pipe_slot>(),
                               static_cast<int>(rest.size()), static_cast<int>(h->P), 0, 1,
                               PIPE_INITIAL, s),
                "initial kernel");
      hip_check(hipStreamSynchronize(s), "sync");
    }
  }
}

}  // namespace

namespace {

// Scratch of the allocation-free resampling entry point
// (dust2_amd_resample_noalloc): the multi-workgroup resampler's partial
// results and cumulative sums for up to kNoAllocMaxParticles weights.  Every
// GPU the library runs on holds its own copy until the process ends; host
// code uses only the per-device address from hipGetSymbolAddress, never
// the array's name.
constexpr size_t kNoAllocMaxParticles = size_t(1) << 20;
constexpr size_t kNoAllocScratchDoubles = 8 + 4 * 128 + kNoAllocMaxParticles;
static __device__ __align__(256) double noalloc_scratch[kNoAllocScratchDoubles];

// The scratch array's device address on the current device, looked up once
// per device (at dust2_amd_init, or the first allocation-free call there).
double* noalloc_scratch_ptr() {
  static std::mutex mu;
  static std::vector<double*> per_device;
  int dev = 0;
  hip_check(hipGetDevice(&dev), "hipGetDevice");
  std::lock_guard<std::mutex> lock(mu);
  if (per_device.size() <= static_cast<size_t>(dev)) per_device.resize(dev + 1, nullptr);
  if (per_device[dev] == nullptr) {
    void* ptr = nullptr;
    hip_check(hipGetSymbolAddress(&ptr, HIP_SYMBOL(noalloc_scratch)), "hipGetSymbolAddress");
    if (ptr == nullptr) throw std::runtime_error("hipGetSymbolAddress returned no address");
    per_device[dev] = static_cast<double*>(ptr);
  }
  return per_device[dev];
}

}  // namespace

extern "C" {

int dust2_amd_init(char* err, size_t errlen) {
  try {
    int n = 0;
    hip_check(hipGetDeviceCount(&n), "hipGetDeviceCount");
    if (n < 1) throw std::runtime_error("no GPU device available");
    tables_for_current_device();
    (void)noalloc_scratch_ptr();
    return 0;
  } catch (const std::exception& e) {
    return fail(e, err, errlen);
  }
}

const char* dust2_amd_version(void) { return "0.1.0"; }

static int filter_alloc_impl(int model, const double* pars, size_t n_groups,
                             size_t n_particles, double time_start, const double* time,
                             size_t n_time, double dt, const double* data, size_t n_groups_data,
                             const uint64_t* seed_state, size_t n_seed_streams, uint64_t seed,
                             int precision, dust2_filter** out, char* err, size_t errlen) {
  std::unique_ptr<dust2_filter> h;
  try {
    check_sizes(n_groups, n_particles);
    check_dt(dt);
    check_time_sequence(time_start, time, n_time);
    if (n_groups_data != 1 && n_groups_data != n_groups) {
      throw std::runtime_error("Expected 'data' to have 1 or " + std::to_string(n_groups) +
                               " groups, but it had " + std::to_string(n_groups_data));
    }
    if (precision != DUST2_FP64 && precision != DUST2_FP32) {
      throw std::runtime_error("'precision' must be DUST2_FP64 or DUST2_FP32");
    }
    const discrete_model_host& m = discrete_model(model);
    if (m.n_data == 0) {
      throw std::runtime_error(std::string("The '") + m.name +
                               "' model does not support 'compare_data'");
    }
    for (size_t g = 0; g < n_groups; ++g) m.check(pars + g * m.n_pars);
    h = std::make_unique<dust2_filter>();
    h->model = model;
    h->S = discrete_state_size(m, pars, n_groups);
    h->n_pars = m.n_pars;
    h->n_data = m.n_data;
    h->kpars = m.kpars;
    h->zero_period = m.zero_period;
    if (discrete_filter_max_particles(model) == 0) h->engine = DUST2_ENGINE_PIPELINE;
    h->fp64 = precision == DUST2_FP64;
    h->G = n_groups;
    h->P = n_particles;
    h->n_time = n_time;
    h->n_groups_data = n_groups_data;
    h->time_start = time_start;
    h->dt = dt;
    h->pars.assign(pars, pars + n_groups * h->n_pars);
    h->times.assign(time, time + n_time);
    const size_t lds = (kScratchDoubles + n_particles) * sizeof(double);
    if (lds > static_cast<size_t>(max_lds_per_block()) ||
        pipe_resample_ppt(static_cast<int>(n_particles)) == 0) {
      // too large for one workgroup's weights: resample across workgroups
      if (mb_resample_blocks(static_cast<int>(n_particles)) > kMbMaxBlocks) {
        throw std::runtime_error("'n_particles' too large: at most " +
                                 std::to_string(static_cast<size_t>(kMbMaxBlocks) * 8192) +
                                 " particles per group");
      }
      h->multi_block = true;
      h->engine = DUST2_ENGINE_PIPELINE;
    }
    h->cfg = discrete_filter_default_config(static_cast<int>(n_particles));
    hip_check(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking), "hipStreamCreate");
    hip_check(hipEventCreateWithFlags(&h->ev_fork, hipEventDisableTiming), "hipEventCreate");
    for (int c = 0; c < dust2_filter::kMaxChains - 1; ++c) {
      hip_check(hipStreamCreateWithFlags(&h->chain_stream[c], hipStreamNonBlocking),
                "hipStreamCreate");
      hip_check(hipEventCreateWithFlags(&h->ev_join[c], hipEventDisableTiming), "hipEventCreate");
    }
    hip_check(hipEventCreate(&h->ev0), "hipEventCreate");
    hip_check(hipEventCreate(&h->ev1), "hipEventCreate");
    const size_t G = n_groups, P = n_particles;
    h->d_times.alloc(n_time * sizeof(double));
    to_device(h->d_times, time, n_time * sizeof(double), h->stream);
    filter_upload_data(h.get(), data);
    h->d_state.alloc(G * h->S * P * h->real_size());
    h->d_rng.alloc(G * 4 * (P + 2) * sizeof(uint64_t));
    {
      dbuf tmp(G * (P + 2) * 32);
      expand_streams(seed_state, n_seed_streams, seed, G * (P + 2), tmp.get<uint64_t>(),
                     h->stream);
      hip_check(launch_rng_aos_to_soa(tmp.get<uint64_t>(), h->d_rng.get<uint64_t>(),
                                      static_cast<int>(G), static_cast<int>(P + 2),
                                      static_cast<int>(P + 2), h->stream),
                "rng layout");
      hip_check(hipStreamSynchronize(h->stream), "sync");
    }
    h->d_ll.alloc(G * sizeof(double));
    h->d_time.alloc(G * sizeof(double));
    h->d_slots.alloc(G * sizeof(filter_slot));
    h->d_errors.alloc(G * P * sizeof(error_record));
    hip_check(hipMemsetAsync(h->d_errors.get(), 0, h->d_errors.bytes(), h->stream), "memset");
    // error counts (chain 0 and 2..4), fallback count (1)
    h->d_counters.alloc(8 * sizeof(uint32_t));
    hip_check(hipMemsetAsync(h->d_counters.get(), 0, h->d_counters.bytes(), h->stream), "memset");
    hip_check(hipMemsetAsync(h->d_state.get(), 0, h->d_state.bytes(), h->stream), "memset");
    hip_check(hipStreamSynchronize(h->stream), "sync");
    *out = h.release();
    return 0;
  } catch (const std::exception& e) {
    return fail(e, err, errlen);
  }
}

int dust2_filter_sir_alloc(const double* pars, size_t n_groups, size_t n_particles,
                           double time_start, const double* time, size_t n_time, double dt,
                           const double* data, size_t n_groups_data,
                           const uint64_t* seed_state, size_t n_seed_streams, uint64_t seed,
                           int precision, dust2_filter** out, char* err, size_t errlen) {
  return filter_alloc_impl(MODEL_SIR, pars, n_groups, n_particles, time_start, time, n_time, dt, data,
                           n_groups_data, seed_state, n_seed_streams, seed, precision, out, err,
                           errlen);
}

int dust2_filter_seir16_alloc(const double* pars, size_t n_groups, size_t n_particles,
                              double time_start, const double* time, size_t n_time, double dt,
                              const double* data, size_t n_groups_data,
                              const uint64_t* seed_state, size_t n_seed_streams, uint64_t seed,
                              int precision, dust2_filter** out, char* err, size_t errlen) {
  return filter_alloc_impl(MODEL_SEIR16, pars, n_groups, n_particles, time_start, time, n_time,
                           dt, data, n_groups_data, seed_state, n_seed_streams, seed, precision,
                           out, err, errlen);
}

int dust2_filter_discrete_alloc(const char* model, const double* pars, size_t n_groups,
                                size_t n_particles, double time_start, const double* time,
                                size_t n_time, double dt, const double* data,
                                size_t n_groups_data, const uint64_t* seed_state,
                                size_t n_seed_streams, uint64_t seed, int precision,
                                dust2_filter** out, char* err, size_t errlen) {
  int id = 0;
  try {
    id = discrete_model_by_name(model);
  } catch (const std::exception& e) {
    return fail(e, err, errlen);
  }
  return filter_alloc_impl(id, pars, n_groups, n_particles, time_start, time, n_time, dt, data,
                           n_groups_data, seed_state, n_seed_streams, seed, precision, out, err,
                           errlen);
}

int dust2_filter_run(dust2_filter* h, const double* initial, int save_trajectories_in,
                     const double* save_snapshots, size_t n_snapshots,
                     const size_t* index_state, size_t n_index_state,
                     const size_t* index_group, size_t n_index, double* ll_out, char* err,
                     size_t errlen) {
  try {
    check_filter(h);
    const auto idx = check_index(index_group, n_index, h->G);
    const auto states = check_state_index(index_state, n_index_state, h->S);
    size_t n_snap = 0;
    const auto slot = snapshot_slots(save_snapshots, n_snapshots, h->times, n_snap);
    // snapshots need the saved states and ancestry of every time, as the
    // reference's filter does (dust2/filter.hpp:47-49)
    const int save_trajectories = (save_trajectories_in || n_snap > 0) ? 1 : 0;
    const size_t G = h->G, P = h->P;
    hipStream_t s = h->stream;
    // trajectories and snapshots become current only when the run completes
    // (dust2/filter.hpp:95-99 sets the flag after the time loop)
    h->traj_valid = false;
    h->ancestry_valid = false;
    if (initial) {
      std::vector<char> tmp(G * h->S * P * h->real_size());
      aos_to_soa(initial, tmp.data(), G, P, h->S, h->fp64);
      to_device(h->d_state, tmp.data(), tmp.size(), s);
    }
    if (save_trajectories && h->d_traj.bytes() == 0) {
      h->d_traj.alloc(h->n_time * G * h->S * P * h->real_size());
    }
    if (save_trajectories && h->d_ancestry.bytes() == 0) {
      h->d_ancestry.alloc(h->n_time * G * P * sizeof(uint32_t));
    }
    std::vector<double> zeros(G, 0.0), t0(G, h->time_start);
    to_device(h->d_ll, zeros.data(), G * sizeof(double), s);
    to_device(h->d_time, t0.data(), G * sizeof(double), s);
    hip_check(hipMemsetAsync(h->d_counters.get(), 0, h->d_counters.bytes(), s), "memset");

    int engine = h->engine;
    if (engine == DUST2_ENGINE_AUTO) {
      // The pipeline spreads each group over several workgroups and, timed
      // over 1-256 groups of 128-2048 SIR particles in FP32 and FP64, is
      // the faster engine except for 256 or more groups of at most 512
      // particles, where one persistent workgroup per group fills the part
      // (by 3-5 %).
      const bool persistent_fits =
          h->beta_n == 0 && static_cast<int>(P) <= discrete_filter_max_particles(h->model);
      engine = (persistent_fits && idx.size() >= 256 && P <= 512) ? DUST2_ENGINE_PERSISTENT
                                                                   : DUST2_ENGINE_PIPELINE;
    }
    if (engine == DUST2_ENGINE_PERSISTENT && h->beta_n > 0) {
      throw std::runtime_error("an interpolated beta runs on the pipeline engine only");
    }
    h->last_engine = engine;
    std::vector<filter_slot> slots;
    if (engine == DUST2_ENGINE_PIPELINE) {
      pipeline_run(h, idx, initial != nullptr, save_trajectories != 0);
    } else {
    std::vector<char> selected(G, 0);
    for (size_t g : idx) selected[g] = 1;
    const uint32_t stride = static_cast<uint32_t>(h->n_groups_data);
    for (size_t g = 0; g < G; ++g) {
      // groups outside index_group are reset (filter.hpp:248-250) but not run
      if (!selected[g] && initial) continue;
      filter_slot sl{};
      discrete_model(h->model).kernel_pars(&h->pars[g * h->n_pars], h->dt, sl.pars);
      sl.times = h->d_times.get<double>();
      const size_t goff = h->n_groups_data == 1 ? 0 : g;
      sl.obs = h->d_obs.get<double>() + goff;
      sl.lfact = h->d_lfact.get<double>() + goff;
      sl.obs_stride = stride;
      sl.n_particles = static_cast<uint32_t>(P);
      sl.t_begin = 0;
      sl.t_end = selected[g] ? static_cast<uint32_t>(h->n_time) : 0;
      sl.flags = (initial ? 0u : SLOT_SET_INITIAL) | (save_trajectories ? SLOT_SAVE_TRAJ : 0u);
      sl.group = static_cast<uint32_t>(g);
      sl.dt = h->dt;
      sl.state = h->d_state.get<char>() + g * h->S * P * h->real_size();
      sl.rng = h->d_rng.get<uint64_t>() + g * 4 * (P + 2);
      sl.ll = h->d_ll.get<double>() + g;
      sl.time = h->d_time.get<double>() + g;
      sl.ancestry = save_trajectories ? h->d_ancestry.get<uint32_t>() + g * h->n_time * P
                                      : nullptr;
      sl.traj = save_trajectories
                    ? h->d_traj.get<char>() + g * h->n_time * h->S * P * h->real_size()
                    : nullptr;
      sl.errors = h->d_errors.get<error_record>() + g * P;
      sl.error_count = h->d_counters.get<uint32_t>();
      sl.n_fallback = h->d_counters.get<uint32_t>() + 1;
      slots.push_back(sl);
    }
    to_device(h->d_slots, slots.data(), slots.size() * sizeof(filter_slot), s);
    hip_check(hipEventRecord(h->ev0, s), "hipEventRecord");
    hip_check(launch_discrete_filter(h->model, h->fp64, h->d_slots.get<filter_slot>(),
                                     static_cast<int>(slots.size()), static_cast<int>(P),
                                     h->cfg, s),
              "filter kernel");
    hip_check(hipEventRecord(h->ev1, s), "hipEventRecord");
    }
    uint32_t counters[8];
    to_host(counters, h->d_counters, sizeof(counters), s);
    float ms = 0;
    hip_check(hipEventElapsedTime(&ms, h->ev0, h->ev1), "hipEventElapsedTime");
    h->last_ms = ms;
    h->last_fallbacks = counters[1];
    h->snap_slot = slot;
    h->n_snap = n_snap;
    h->traj_states = states;
    h->last_index = idx;
    // a new run forgets the chosen particles of its groups (filter.hpp:162-175)
    if (h->select_valid.size() != G) h->select_valid.assign(G, 0);
    for (size_t g : idx) h->select_valid[g] = 0;
    if (counters[0] + counters[2] + counters[3] + counters[4] > 0) {
      std::vector<error_record> rec(G * P);
      to_host(rec.data(), h->d_errors, rec.size() * sizeof(error_record), s);
      hip_check(hipMemsetAsync(h->d_errors.get(), 0, h->d_errors.bytes(), s), "memset");
      hip_check(hipStreamSynchronize(s), "sync");
      // the run stops at the first interval with errors
      // (dust2/discrete/system.hpp:87): independent launch chains may have
      // gone on to fail later, so keep only the earliest interval's errors
      uint32_t first = UINT32_MAX;
      for (const auto& r : rec) {
        if (r.code != ERR_NONE) first = std::min(first, r.interval);
      }
      size_t count = 0;
      for (auto& r : rec) {
        if (r.code != ERR_NONE && r.interval != first) r.code = ERR_NONE;
        if (r.code != ERR_NONE) ++count;
      }
      throw std::runtime_error(error_report(rec, count));
    }
    h->traj_valid = save_trajectories_in != 0;
    h->ancestry_valid = save_trajectories != 0;
    std::vector<double> ll(G);
    to_host(ll.data(), h->d_ll, G * sizeof(double), s);
    for (size_t i = 0; i < idx.size(); ++i) ll_out[i] = ll[idx[i]];
    return 0;
  } catch (const std::exception& e) {
    return fail(e, err, errlen);
  }
}

int dust2_filter_update_pars(dust2_filter* h, const double* pars, const size_t* index_group,
                                 size_t n_index, char* err, size_t errlen) {
  try {
    check_filter(h);
    const auto idx = check_index(index_group, n_index, h->G);
    for (size_t i = 0; i < idx.size(); ++i) {
      const double* row = pars + i * h->n_pars;
      double* dst = &h->pars[idx[i] * h->n_pars];
      const discrete_model_host& m = discrete_model(h->model);
      m.check(row);
      for (size_t k : m.updatable) dst[k] = row[k];
    }
    return 0;
  } catch (const std::exception& e) {
    return fail(e, err, errlen);
  }
}

int dust2_filter_last_state(dust2_filter* h, double* out, char* err, size_t errlen) {
  try {
    check_filter(h);
    std::vector<char> tmp(h->d_state.bytes());
    to_host(tmp.data(), h->d_state, tmp.size(), h->stream);
    soa_to_aos(tmp.data(), out, h->G, h->P, h->S, h->fp64);
    return 0;
  } catch (const std::exception& e) {
    return fail(e, err, errlen);
  }
}

namespace {
// Groups of the last run as u32 on the device; chosen particles for them
// (drawn once per run, filter.hpp:130-138) when `select` is set.
const uint32_t* lineage_groups(dust2_filter* h, bool select) {
  std::vector<uint32_t> gs(h->last_index.begin(), h->last_index.end());
  if (h->d_groups.bytes() < h->G * sizeof(uint32_t)) h->d_groups.alloc(h->G * sizeof(uint32_t));
  to_device(h->d_groups, gs.data(), gs.size() * sizeof(uint32_t), h->stream);
  if (select) {
    if (h->d_select.bytes() < h->G * sizeof(uint32_t)) {
      h->d_select.alloc(h->G * sizeof(uint32_t));
    }
    std::vector<uint32_t> todo;
    for (size_t g : h->last_index) {
      if (!h->select_valid[g]) todo.push_back(static_cast<uint32_t>(g));
    }
    if (!todo.empty()) {
      dbuf d_todo;
      d_todo.alloc(todo.size() * sizeof(uint32_t));
      to_device(d_todo, todo.data(), todo.size() * sizeof(uint32_t), h->stream);
      hip_check(launch_select_particle(h->d_rng.get<uint64_t>(), d_todo.get<uint32_t>(),
                                       static_cast<int>(todo.size()), static_cast<int>(h->P),
                                       4 * (h->P + 2), 1, h->P + 2,
                                       h->d_select.get<uint32_t>(), h->stream),
                "select particle kernel");
      hip_check(hipStreamSynchronize(h->stream), "sync");
      for (uint32_t g : todo) h->select_valid[g] = 1;
    }
  }
  return h->d_groups.get<uint32_t>();
}
}  // namespace

int dust2_filter_last_trajectories(dust2_filter* h, int select_random_particle, double* out,
                                   char* err, size_t errlen) {
  try {
    check_filter(h);
    if (!h->traj_valid) {
      throw std::runtime_error("Trajectories are not current");
    }
    const bool select = select_random_particle != 0;
    const uint32_t* groups = lineage_groups(h, select);
    const size_t n_out = h->last_index.size();
    const size_t p_out = select ? 1 : h->P;
    const size_t s_out = h->traj_states.empty() ? h->S : h->traj_states.size();
    const size_t bytes = h->n_time * n_out * p_out * s_out * sizeof(double);
    if (h->d_traj_out.bytes() < bytes) h->d_traj_out.alloc(bytes);
    if (!h->traj_states.empty()) {
      if (h->d_traj_states.bytes() < h->S * sizeof(uint32_t)) {
        h->d_traj_states.alloc(h->S * sizeof(uint32_t));
      }
      to_device(h->d_traj_states, h->traj_states.data(), s_out * sizeof(uint32_t), h->stream);
    }
    hip_check(launch_lineage(h->d_traj.get(), h->d_ancestry.get<uint32_t>(), groups,
                             static_cast<int>(n_out), static_cast<int>(h->n_time),
                             static_cast<int>(h->S), static_cast<int>(h->P),
                             select ? h->d_select.get<uint32_t>() : nullptr, nullptr,
                             h->traj_states.empty() ? nullptr : h->d_traj_states.get<uint32_t>(),
                             static_cast<int>(s_out), h->fp64, h->d_traj_out.get<double>(),
                             h->stream),
              "lineage kernel");
    to_host(out, h->d_traj_out, bytes, h->stream);
    return 0;
  } catch (const std::exception& e) {
    return fail(e, err, errlen);
  }
}

int dust2_filter_last_trajectories_raw(dust2_filter* h, double* out, char* err,
                                       size_t errlen) {
  try {
    check_filter(h);
    if (!h->traj_valid) {
      throw std::runtime_error("Trajectories are not current");
    }
    // device: [G][n_time][S][P]
    const size_t G = h->G, P = h->P, T = h->n_time;
    std::vector<char> tmp(h->d_traj.bytes());
    to_host(tmp.data(), h->d_traj, tmp.size(), h->stream);
    for (size_t g = 0; g < G; ++g) {
      for (size_t t = 0; t < T; ++t) {
        for (size_t c = 0; c < h->S; ++c) {
          for (size_t p = 0; p < P; ++p) {
            const size_t o = ((g * T + t) * h->S + c) * P + p;
            out[((t * G + g) * P + p) * h->S + c] =
                h->fp64 ? reinterpret_cast<const double*>(tmp.data())[o]
                        : reinterpret_cast<const float*>(tmp.data())[o];
          }
        }
      }
    }
    return 0;
  } catch (const std::exception& e) {
    return fail(e, err, errlen);
  }
}

int dust2_filter_last_snapshots(dust2_filter* h, int select_random_particle, double* out,
                                char* err, size_t errlen) {
  try {
    check_filter(h);
    if (!h->ancestry_valid || h->n_snap == 0) {
      throw std::runtime_error("Snapshots are not current");
    }
    const bool select = select_random_particle != 0;
    const uint32_t* groups = lineage_groups(h, select);
    const size_t n_out = h->last_index.size();
    const size_t p_out = select ? 1 : h->P;
    const size_t bytes = h->n_snap * n_out * p_out * h->S * sizeof(double);
    if (h->d_traj_out.bytes() < bytes) h->d_traj_out.alloc(bytes);
    dbuf d_slot(h->n_time * sizeof(int32_t));
    to_device(d_slot, h->snap_slot.data(), h->n_time * sizeof(int32_t), h->stream);
    hip_check(launch_lineage(h->d_traj.get(), h->d_ancestry.get<uint32_t>(), groups,
                             static_cast<int>(n_out), static_cast<int>(h->n_time),
                             static_cast<int>(h->S), static_cast<int>(h->P),
                             select ? h->d_select.get<uint32_t>() : nullptr,
                             d_slot.get<int32_t>(), nullptr, 0, h->fp64,
                             h->d_traj_out.get<double>(), h->stream),
              "lineage kernel");
    to_host(out, h->d_traj_out, bytes, h->stream);
    return 0;
  } catch (const std::exception& e) {
    return fail(e, err, errlen);
  }
}

int dust2_filter_last_state_selected(dust2_filter* h, double* out, char* err, size_t errlen) {
  try {
    check_filter(h);
    if (h->last_index.empty()) {
      throw std::runtime_error("State is not current");
    }
    lineage_groups(h, true);
    std::vector<uint32_t> sel(h->G);
    to_host(sel.data(), h->d_select, h->G * sizeof(uint32_t), h->stream);
    std::vector<char> tmp(h->d_state.bytes());
    to_host(tmp.data(), h->d_state, tmp.size(), h->stream);
    const size_t P = h->P, S = h->S;
    for (size_t i = 0; i < h->last_index.size(); ++i) {
      const size_t g = h->last_index[i];
      for (size_t c = 0; c < S; ++c) {
        const size_t o = (g * S + c) * P + sel[g];
        out[i * S + c] = h->fp64 ? reinterpret_cast<const double*>(tmp.data())[o]
                                 : reinterpret_cast<const float*>(tmp.data())[o];
      }
    }
    return 0;
  } catch (const std::exception& e) {
    return fail(e, err, errlen);
  }
}

int dust2_filter_last_ancestry(dust2_filter* h, uint32_t* out, char* err, size_t errlen) {
  try {
    check_filter(h);
    if (!h->ancestry_valid) {
      throw std::runtime_error("Ancestry is not current; run with 'save_trajectories'");
    }
    const size_t G = h->G, P = h->P, T = h->n_time;
    std::vector<uint32_t> tmp(G * T * P);
    to_host(tmp.data(), h->d_ancestry, tmp.size() * sizeof(uint32_t), h->stream);
    for (size_t g = 0; g < G; ++g) {
      for (size_t t = 0; t < T; ++t) {
        std::memcpy(out + (t * G + g) * P, tmp.data() + (g * T + t) * P, P * sizeof(uint32_t));
      }
    }
    return 0;
  } catch (const std::exception& e) {
    return fail(e, err, errlen);
  }
}

int dust2_filter_rng_state(dust2_filter* h, uint64_t* out, char* err, size_t errlen) {
  try {
    check_filter(h);
    const size_t n = h->G * (h->P + 2);
    dbuf tmp(n * 32);
    hip_check(launch_rng_soa_to_aos(h->d_rng.get<uint64_t>(), tmp.get<uint64_t>(),
                                    static_cast<int>(h->G), static_cast<int>(h->P + 2),
                                    static_cast<int>(h->P + 2), h->stream),
              "rng layout");
    to_host(out, tmp, n * 32, h->stream);
    return 0;
  } catch (const std::exception& e) {
    return fail(e, err, errlen);
  }
}

int dust2_filter_set_rng_state(dust2_filter* h, const uint64_t* state, char* err,
                                   size_t errlen) {
  try {
    check_filter(h);
    const size_t n = h->G * (h->P + 2);
    dbuf tmp(n * 32);
    to_device(tmp, state, n * 32, h->stream);
    hip_check(launch_rng_aos_to_soa(tmp.get<uint64_t>(), h->d_rng.get<uint64_t>(),
                                    static_cast<int>(h->G), static_cast<int>(h->P + 2),
                                    static_cast<int>(h->P + 2), h->stream),
              "rng layout");
    hip_check(hipStreamSynchronize(h->stream), "sync");
    return 0;
  } catch (const std::exception& e) {
    return fail(e, err, errlen);
  }
}

int dust2_filter_set_launch(dust2_filter* h, int block, int ppt, char* err, size_t errlen) {
  try {
    check_filter(h);
    const launch_config c{block, ppt};
    if (!discrete_filter_config_valid(c)) {
      throw std::runtime_error("unsupported launch shape");
    }
    if (static_cast<size_t>(block) * ppt < h->P) {
      throw std::runtime_error("launch shape holds fewer than n_particles particles");
    }
    h->cfg = c;
    return 0;
  } catch (const std::exception& e) {
    return fail(e, err, errlen);
  }
}

int dust2_filter_last_stats(dust2_filter* h, double* out, char* err, size_t errlen) {
  try {
    check_filter(h);
    out[0] = h->last_fallbacks;
    out[1] = h->last_ms;
    out[2] = h->last_engine;
    out[3] = h->last_advance_ms;
    out[4] = h->last_resample_ms;
    out[5] = h->cfg.block;
    out[6] = h->cfg.ppt;
    return 0;
  } catch (const std::exception& e) {
    return fail(e, err, errlen);
  }
}

int dust2_filter_set_beta_schedule(dust2_filter* h, int kind, const double* t, size_t n_t,
                                   const double* y, size_t n_groups_y, char* err,
                                   size_t errlen) {
  try {
    check_filter(h);
    if (kind < 0) {
      h->beta_kind = -1;
      h->beta_n = 0;
      h->beta_groups = 0;
      return 0;
    }
    if (h->model != MODEL_SIR) {
      throw std::runtime_error("an interpolated beta is a SIR model input");
    }
    if (kind != INTERP_CONSTANT && kind != INTERP_LINEAR && kind != INTERP_SPLINE) {
      throw std::runtime_error("unknown interpolation kind");
    }
    if (n_groups_y != 1 && n_groups_y != h->G) {
      throw std::runtime_error("'beta_y' must be shared by all groups or given for each");
    }
    interp_validate(t, n_t, n_t, "beta_t", "beta_y");
    if (n_t == 0 || (kind != INTERP_CONSTANT && n_t < 2)) {
      throw std::runtime_error("'beta_t' needs at least " +
                               std::string(kind == INTERP_CONSTANT ? "one time" : "two times"));
    }
    // slopes per group: y is [n_groups_y][n_t], one series each
    std::vector<double> k(n_groups_y * n_t, 0.0);
    if (kind == INTERP_SPLINE) {
      for (size_t g = 0; g < n_groups_y; ++g) {
        const std::vector<double> kg = interp_spline_slopes(t, y + g * n_t, n_t, 1);
        std::copy(kg.begin(), kg.end(), k.begin() + g * n_t);
      }
    }
    hipStream_t s = h->stream;
    h->d_beta_t.alloc(n_t * sizeof(double));
    h->d_beta_y.alloc(n_groups_y * n_t * sizeof(double));
    h->d_beta_k.alloc(n_groups_y * n_t * sizeof(double));
    to_device(h->d_beta_t, t, n_t * sizeof(double), s);
    to_device(h->d_beta_y, y, n_groups_y * n_t * sizeof(double), s);
    to_device(h->d_beta_k, k.data(), k.size() * sizeof(double), s);
    hip_check(hipStreamSynchronize(s), "sync");
    h->beta_kind = kind;
    h->beta_n = n_t;
    h->beta_groups = n_groups_y;
    return 0;
  } catch (const std::exception& e) {
    return fail(e, err, errlen);
  }
}

int dust2_filter_set_pipeline_chains(dust2_filter* h, int chains, char* err, size_t errlen) {
  try {
    check_filter(h);
    if (chains < 1 || chains > dust2_filter::kMaxChains) {
      throw std::runtime_error("'chains' must be between 1 and " +
                               std::to_string(dust2_filter::kMaxChains));
    }
    h->pipe_chains = chains;
    return 0;
  } catch (const std::exception& e) {
    return fail(e, err, errlen);
  }
}

int dust2_filter_set_pipeline_resample(dust2_filter* h, int block, char* err, size_t errlen) {
  try {
    check_filter(h);
    if (block != 1024 && block != 512 && block != 256) {
      throw std::runtime_error("'block' must be 1024, 512 or 256");
    }
    if (!h->multi_block && pipe_resample_ppt(static_cast<int>(h->P), block) == 0) {
      throw std::runtime_error("a resample workgroup of " + std::to_string(block) +
                               " threads does not hold " + std::to_string(h->P) +
                               " particles");
    }
    h->pipe_resample_block = block;
    return 0;
  } catch (const std::exception& e) {
    return fail(e, err, errlen);
  }
}

int dust2_filter_set_engine(dust2_filter* h, int engine, int use_graph, int profile,
                                char* err, size_t errlen) {
  try {
    check_filter(h);
    if (engine != DUST2_ENGINE_AUTO && engine != DUST2_ENGINE_PERSISTENT &&
        engine != DUST2_ENGINE_PIPELINE) {
      throw std::runtime_error("unknown engine");
    }
    if (engine == DUST2_ENGINE_PERSISTENT && discrete_filter_max_particles(h->model) == 0) {
      throw std::runtime_error("the persistent engine runs the sir model only");
    }
    if (engine == DUST2_ENGINE_PERSISTENT &&
        static_cast<int>(h->P) > discrete_filter_max_particles(h->model)) {
      throw std::runtime_error("the persistent engine holds at most " +
                               std::to_string(discrete_filter_max_particles(h->model)) +
                               " particles per group");
    }
    h->engine = engine;
    h->use_graph = use_graph != 0;
    h->profile = profile != 0;
    return 0;
  } catch (const std::exception& e) {
    return fail(e, err, errlen);
  }
}

int dust2_filter_autotune(dust2_filter* h, const size_t* index_group, size_t n_index,
                          double* report, size_t report_rows, size_t* n_report, char* err,
                          size_t errlen) {
  try {
    check_filter(h);
    const auto idx = check_index(index_group, n_index, h->G);
    const int P = static_cast<int>(h->P);
    // candidates: the persistent kernel at every launch shape that holds the
    // group without more than doubling it, and the interval pipeline
    struct candidate {
      int engine;
      launch_config cfg;
      int chains;
      int rblock;
    };
    std::vector<candidate> cand;
    if (h->beta_n == 0 && P <= discrete_filter_max_particles(h->model)) {
      static const int shapes[][2] = {{64, 1},   {64, 2},   {64, 4},   {128, 2}, {128, 4},
                                      {256, 2},  {256, 4},  {256, 8},  {512, 2}, {512, 4},
                                      {512, 8},  {1024, 2}, {1024, 4}, {1024, 8}};
      for (const auto& sh : shapes) {
        const int cap = sh[0] * sh[1];
        if (cap >= P && cap < 2 * P + 64) {
          cand.push_back({DUST2_ENGINE_PERSISTENT, launch_config{sh[0], sh[1]}, 1, 1024});
        }
      }
    }
    for (int chains = 1; chains <= dust2_filter::kMaxChains; ++chains) {
      if (chains > 1 && idx.size() < static_cast<size_t>(chains)) continue;
      for (const int rblock : {1024, 512, 256}) {
        if (rblock != 1024 && (h->multi_block || pipe_resample_ppt(P, rblock) == 0)) continue;
        cand.push_back({DUST2_ENGINE_PIPELINE, h->cfg, chains, rblock});
      }
    }
    // tuning runs must leave the object as they found it: the particle and
    // filter streams and the state are restored after every run
    hipStream_t s = h->stream;
    dbuf rng0(h->d_rng.bytes()), state0(h->d_state.bytes());
    hip_check(hipMemcpyAsync(rng0.get(), h->d_rng.get(), h->d_rng.bytes(),
                             hipMemcpyDeviceToDevice, s), "copy");
    hip_check(hipMemcpyAsync(state0.get(), h->d_state.get(), h->d_state.bytes(),
                             hipMemcpyDeviceToDevice, s), "copy");
    const int engine0 = h->engine;
    const launch_config cfg0 = h->cfg;
    const bool graph0 = h->use_graph, profile0 = h->profile;
    const int chains0 = h->pipe_chains;
    const int rblock0 = h->pipe_resample_block;
    std::vector<double> ll(idx.size());
    std::vector<double> ms(cand.size(), 0.0);
    std::string failure;
    for (size_t i = 0; i < cand.size() && failure.empty(); ++i) {
      h->engine = cand[i].engine;
      h->cfg = cand[i].cfg;
      h->pipe_chains = cand[i].chains;
      h->pipe_resample_block = cand[i].rblock;
      h->use_graph = true;
      h->profile = false;
      // a cold run (graph capture), then the faster of two warm runs
      for (int rep = 0; rep < 3; ++rep) {
        char buf[4096];
        const int rc = dust2_filter_run(h, nullptr, 0, nullptr, 0, nullptr, 0, idx.data(),
                                        idx.size(), ll.data(), buf, sizeof buf);
        hip_check(hipMemcpyAsync(h->d_rng.get(), rng0.get(), h->d_rng.bytes(),
                                 hipMemcpyDeviceToDevice, s), "copy");
        if (rc != 0) {
          failure = buf;
          break;
        }
        if (rep == 1 || (rep == 2 && h->last_ms < ms[i])) ms[i] = h->last_ms;
      }
    }
    hip_check(hipMemcpyAsync(h->d_state.get(), state0.get(), h->d_state.bytes(),
                             hipMemcpyDeviceToDevice, s), "copy");
    hip_check(hipStreamSynchronize(s), "sync");
    h->use_graph = graph0;
    h->profile = profile0;
    h->traj_valid = false;
    h->ancestry_valid = false;
    if (!failure.empty()) {
      h->engine = engine0;
      h->cfg = cfg0;
      h->pipe_chains = chains0;
      h->pipe_resample_block = rblock0;
      throw std::runtime_error(failure);
    }
    size_t best = 0;
    for (size_t i = 1; i < cand.size(); ++i) {
      if (ms[i] < ms[best]) best = i;
    }
    h->engine = cand[best].engine;
    h->cfg = h->engine == DUST2_ENGINE_PERSISTENT ? cand[best].cfg : cfg0;
    h->pipe_chains = h->engine == DUST2_ENGINE_PIPELINE ? cand[best].chains : chains0;
    h->pipe_resample_block =
        h->engine == DUST2_ENGINE_PIPELINE ? cand[best].rblock : rblock0;
    const size_t n = std::min(report_rows, cand.size());
    for (size_t i = 0; report && i < n; ++i) {
      const bool pers = cand[i].engine == DUST2_ENGINE_PERSISTENT;
      report[4 * i + 0] = cand[i].engine;
      report[4 * i + 1] = pers ? cand[i].cfg.block : cand[i].rblock;
      report[4 * i + 2] = pers ? cand[i].cfg.ppt : cand[i].chains;
      report[4 * i + 3] = ms[i];
    }
    if (n_report) *n_report = cand.size();
    return 0;
  } catch (const std::exception& e) {
    return fail(e, err, errlen);
  }
}

void dust2_filter_free(dust2_filter* h) {
  if (handle_ok(h, TAG_FILTER)) delete h;
}

}  // extern "C"

// ============================================================================
// System
// ============================================================================
struct dust2_system {
  uint32_t tag = TAG_SYSTEM;
  int model = MODEL_SIR;     // discrete_model_id of models.hpp
  size_t S = 5;              // states per particle
  bool fp64 = true;
  bool deterministic = false;
  size_t G = 0, P = 0;
  double time = 0, dt = 1;
  std::vector<double> pars;  // G x n_pars host rows
  hipStream_t stream = nullptr;
  dbuf d_state, d_tmp, d_rng, d_pars, d_errors, d_counters, d_obs, d_lfact, d_out, d_idx;
  bool errors_pending = false;
  size_t real_size() const { return fp64 ? 8 : 4; }
  ~dust2_system() {
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace {

void check_system(dust2_system* h, const char* action) {
  check_handle(h, TAG_SYSTEM, "system");
  if (h->errors_pending) {
    throw std::runtime_error(std::string("Can't currently ") + action +
                             " this system: errors are pending");
  }
}

void system_upload_pars(dust2_system* h) {
  const discrete_model_host& m = discrete_model(h->model);
  std::vector<double> kp(h->G * m.kpars);
  for (size_t g = 0; g < h->G; ++g) m.kernel_pars(&h->pars[g * m.n_pars], h->dt, &kp[g * m.kpars]);
  if (h->d_pars.bytes() < kp.size() * sizeof(double)) h->d_pars.alloc(kp.size() * sizeof(double));
  to_device(h->d_pars, kp.data(), kp.size() * sizeof(double), h->stream);
}

discrete_run_args system_args(dust2_system* h) {
  discrete_run_args a{};
  a.state = h->d_state.get();
  a.rng = h->d_rng.get<uint64_t>();
  a.kpars = h->d_pars.get<double>();
  a.kstride = static_cast<int>(discrete_model(h->model).kpars);
  a.G = static_cast<int>(h->G);
  a.P = static_cast<int>(h->P);
  a.time = h->time;
  a.dt = h->dt;
  a.flags = h->deterministic ? SLOT_DETERMINISTIC : 0u;
  a.errors = h->d_errors.get<error_record>();
  a.error_count = h->d_counters.get<uint32_t>();
  return a;
}

void system_run(dust2_system* h, double time, bool set_initial) {
  const double diff = std::max(0.0, time - h->time);
  const size_t n_steps = static_cast<size_t>(std::round(diff / h->dt));
  hipStream_t s = h->stream;
  hip_check(hipMemsetAsync(h->d_counters.get(), 0, h->d_counters.bytes(), s), "memset");
  discrete_run_args a = system_args(h);
  a.n_steps = static_cast<int>(n_steps);
  if (set_initial) a.flags |= SLOT_SET_INITIAL;
  hip_check(launch_discrete_run(h->model, h->fp64, a, s), "run_to_time kernel");
  uint32_t count = 0;
  to_host(&count, h->d_counters, sizeof(count), s);
  if (count == 0) {
    // on error the reference throws before advancing time
    // (dust2/discrete/system.hpp:87-91)
    h->time = h->time + n_steps * h->dt;
  } else {
    std::vector<error_record> rec(h->G * h->P);
    to_host(rec.data(), h->d_errors, rec.size() * sizeof(error_record), s);
    hip_check(hipMemsetAsync(h->d_errors.get(), 0, h->d_errors.bytes(), s), "memset");
    hip_check(hipStreamSynchronize(s), "sync");
    h->errors_pending = true;
    throw std::runtime_error(error_report(rec, count));
  }
}

}  // namespace

extern "C" {

int dust2_system_discrete_alloc(const char* model, const double* pars, size_t n_groups,
                                size_t n_particles, double time, double dt,
                                const uint64_t* seed_state, size_t n_seed_streams, uint64_t seed,
                                int deterministic, int precision, dust2_system** out, char* err,
                                size_t errlen) {
  try {
    const discrete_model_host& m = discrete_model(discrete_model_by_name(model));
    check_sizes(n_groups, n_particles);
    check_dt(dt);
    if (precision != DUST2_FP64 && precision != DUST2_FP32) {
      throw std::runtime_error("'precision' must be DUST2_FP64 or DUST2_FP32");
    }
    for (size_t g = 0; g < n_groups; ++g) m.check(pars + g * m.n_pars);
    auto h = std::make_unique<dust2_system>();
    h->model = m.id;
    h->S = discrete_state_size(m, pars, n_groups);
    h->fp64 = precision == DUST2_FP64;
    h->deterministic = deterministic != 0;
    h->G = n_groups;
    h->P = n_particles;
    h->time = time;
    h->dt = dt;
    h->pars.assign(pars, pars + n_groups * m.n_pars);
    hip_check(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking), "hipStreamCreate");
    const size_t G = n_groups, P = n_particles;
    h->d_state.alloc(G * h->S * P * h->real_size());
    h->d_tmp.alloc(G * h->S * P * h->real_size());
    h->d_rng.alloc(G * 4 * P * sizeof(uint64_t));
    h->d_errors.alloc(G * P * sizeof(error_record));
    h->d_counters.alloc(4 * sizeof(uint32_t));
    hip_check(hipMemsetAsync(h->d_state.get(), 0, h->d_state.bytes(), h->stream), "memset");
    hip_check(hipMemsetAsync(h->d_errors.get(), 0, h->d_errors.bytes(), h->stream), "memset");
    {
      dbuf tmp(G * P * 32);
      expand_streams(seed_state, n_seed_streams, seed, G * P, tmp.get<uint64_t>(), h->stream);
      hip_check(launch_rng_aos_to_soa(tmp.get<uint64_t>(), h->d_rng.get<uint64_t>(),
                                      static_cast<int>(G), static_cast<int>(P),
                                      static_cast<int>(P), h->stream),
                "rng layout");
      hip_check(hipStreamSynchronize(h->stream), "sync");
    }
    system_upload_pars(h.get());
    hip_check(hipStreamSynchronize(h->stream), "sync");
    *out = h.release();
    return 0;
  } catch (const std::exception& e) {
    return fail(e, err, errlen);
  }
}

int dust2_system_sir_alloc(const double* pars, size_t n_groups, size_t n_particles, double time,
                           double dt, const uint64_t* seed_state, size_t n_seed_streams,
                           uint64_t seed, int deterministic, int precision, dust2_system** out,
                           char* err, size_t errlen) {
  return dust2_system_discrete_alloc("sir", pars, n_groups, n_particles, time, dt, seed_state,
                                     n_seed_streams, seed, deterministic, precision, out, err,
                                     errlen);
}

int dust2_system_discrete_run_to_time(dust2_system* h, double time, char* err, size_t errlen) {
  try {
    check_system(h, "run");
    if (time < h->time) {
      // dust2/r/system.hpp:24-27
      char buf[256];
      std::snprintf(buf, sizeof buf, "Can't run to time %f, system already at time %f", time,
                    h->time);
      throw std::runtime_error(buf);
    }
    system_run(h, time, false);
    return 0;
  } catch (const std::exception& e) {
    return fail(e, err, errlen);
  }
}

int dust2_system_discrete_simulate(dust2_system* h, const double* times, size_t n_times,
                              double* out, char* err, size_t errlen) {
  try {
    check_system(h, "simulate");
    const size_t block = h->G * h->P * h->S;
    for (size_t i = 0; i < n_times; ++i) {
      system_run(h, times[i], false);
      std::vector<char> tmp(h->d_state.bytes());
      to_host(tmp.data(), h->d_state, tmp.size(), h->stream);
      soa_to_aos(tmp.data(), out + i * block, h->G, h->P, h->S, h->fp64);
    }
    return 0;
  } catch (const std::exception& e) {
    return fail(e, err, errlen);
  }
}

int dust2_system_discrete_state(dust2_system* h, double* out, char* err, size_t errlen) {
  try {
    check_system(h, "get state from");
    std::vector<char> tmp(h->d_state.bytes());
    to_host(tmp.data(), h->d_state, tmp.size(), h->stream);
    soa_to_aos(tmp.data(), out, h->G, h->P, h->S, h->fp64);
    return 0;
  } catch (const std::exception& e) {
    return fail(e, err, errlen);
  }
}

int dust2_system_discrete_set_state(dust2_system* h, const double* state, char* err, size_t errlen) {
  try {
    check_handle(h, TAG_SYSTEM, "system");  // allowed while errors are pending
    std::vector<char> tmp(h->d_state.bytes());
    aos_to_soa(state, tmp.data(), h->G, h->P, h->S, h->fp64);
    to_device(h->d_state, tmp.data(), tmp.size(), h->stream);
    hip_check(hipStreamSynchronize(h->stream), "sync");
    h->errors_pending = false;
    return 0;
  } catch (const std::exception& e) {
    return fail(e, err, errlen);
  }
}

int dust2_system_discrete_set_state_initial(dust2_system* h, char* err, size_t errlen) {
  try {
    check_handle(h, TAG_SYSTEM, "system");  // allowed while errors are pending
    h->errors_pending = false;
    // a zero-step run from the initial state writes initial(), drawing
    // from the particles' streams where the model's initial() does
    // (walk.cpp random_initial)
    system_run(h, h->time, true);
    return 0;
  } catch (const std::exception& e) {
    return fail(e, err, errlen);
  }
}

int dust2_system_discrete_time(dust2_system* h, double* out, char* err, size_t errlen) {
  try {
    check_system(h, "get time from");
    *out = h->time;
    return 0;
  } catch (const std::exception& e) {
    return fail(e, err, errlen);
  }
}

int dust2_system_discrete_set_time(dust2_system* h, double time, char* err, size_t errlen) {
  try {
    check_system(h, "set time for");
    h->time = time;
    return 0;
  } catch (const std::exception& e) {
    return fail(e, err, errlen);
  }
}

int dust2_system_discrete_update_pars(dust2_system* h, const double* pars, char* err, size_t errlen) {
  try {
    check_handle(h, TAG_SYSTEM, "system");  // allowed while errors are pending
    const discrete_model_host& m = discrete_model(h->model);
    for (size_t g = 0; g < h->G; ++g) m.check(pars + g * m.n_pars);
    (void)discrete_state_size(m, pars, h->G);
    if (m.n_state(pars) != h->S) {
      throw std::runtime_error("'pars' would change the state length of the system");
    }
    for (size_t g = 0; g < h->G; ++g) {
      const double* row = pars + g * m.n_pars;
      for (size_t k : m.updatable) h->pars[g * m.n_pars + k] = row[k];
    }
    system_upload_pars(h);
    hip_check(hipStreamSynchronize(h->stream), "sync");
    return 0;
  } catch (const std::exception& e) {
    return fail(e, err, errlen);
  }
}

int dust2_system_discrete_compare_data(dust2_system* h, const double* data, size_t n_groups_data,
                                  double* out, char* err, size_t errlen) {
  try {
    check_system(h, "compare data for");
    const size_t nd = discrete_model(h->model).n_data;
    if (nd == 0) {
      throw std::runtime_error(std::string("The '") + discrete_model(h->model).name +
                               "' model does not support 'compare_data'");
    }
    if (n_groups_data != 1 && n_groups_data != h->G) {
      throw std::runtime_error("Expected 'data' to have length " + std::to_string(h->G) +
                               ", but it had length " + std::to_string(n_groups_data));
    }
    const size_t n = n_groups_data * nd;
    std::vector<double> lf(n);
    for (size_t i = 0; i < n; ++i) {
      lf[i] = std::isnan(data[i]) ? 0.0 : std::lgamma(data[i] + 1);
    }
    hipStream_t s = h->stream;
    h->d_obs.alloc(n * sizeof(double));
    h->d_lfact.alloc(n * sizeof(double));
    h->d_out.alloc(h->G * h->P * sizeof(double));
    to_device(h->d_obs, data, n * sizeof(double), s);
    to_device(h->d_lfact, lf.data(), n * sizeof(double), s);
    hip_check(launch_discrete_compare(h->model, h->fp64, system_args(h), h->d_obs.get<double>(),
                                      h->d_lfact.get<double>(),
                                      n_groups_data == 1 ? 0 : static_cast<int>(nd),
                                      h->d_out.get<double>(), s),
              "compare kernel");
    to_host(out, h->d_out, h->G * h->P * sizeof(double), s);
    return 0;
  } catch (const std::exception& e) {
    return fail(e, err, errlen);
  }
}

int dust2_system_discrete_reorder(dust2_system* h, const uint32_t* index, char* err, size_t errlen) {
  try {
    check_system(h, "reorder");
    for (size_t i = 0; i < h->G * h->P; ++i) {
      if (index[i] >= h->P) throw std::runtime_error("Invalid index for reorder");
    }
    hipStream_t s = h->stream;
    h->d_idx.alloc(h->G * h->P * sizeof(uint32_t));
    to_device(h->d_idx, index, h->G * h->P * sizeof(uint32_t), s);
    hip_check(launch_reorder(h->d_state.get(), h->d_tmp.get(), h->d_idx.get<uint32_t>(),
                             static_cast<int>(h->G), h->S, static_cast<int>(h->P), h->fp64,
                             s),
              "reorder kernel");
    hip_check(hipStreamSynchronize(s), "sync");
    h->d_state.swap(h->d_tmp);
    return 0;
  } catch (const std::exception& e) {
    return fail(e, err, errlen);
  }
}

int dust2_system_discrete_rng_state(dust2_system* h, uint64_t* out, char* err, size_t errlen) {
  try {
    check_handle(h, TAG_SYSTEM, "system");  // allowed while errors are pending
    const size_t n = h->G * h->P;
    dbuf tmp(n * 32);
    hip_check(launch_rng_soa_to_aos(h->d_rng.get<uint64_t>(), tmp.get<uint64_t>(),
                                    static_cast<int>(h->G), static_cast<int>(h->P),
                                    static_cast<int>(h->P), h->stream),
              "rng layout");
    to_host(out, tmp, n * 32, h->stream);
    return 0;
  } catch (const std::exception& e) {
    return fail(e, err, errlen);
  }
}

int dust2_system_discrete_set_rng_state(dust2_system* h, const uint64_t* state, char* err,
                                   size_t errlen) {
  try {
    check_handle(h, TAG_SYSTEM, "system");  // allowed while errors are pending
    const size_t n = h->G * h->P;
    dbuf tmp(n * 32);
    to_device(tmp, state, n * 32, h->stream);
    hip_check(launch_rng_aos_to_soa(tmp.get<uint64_t>(), h->d_rng.get<uint64_t>(),
                                    static_cast<int>(h->G), static_cast<int>(h->P),
                                    static_cast<int>(h->P), h->stream),
              "rng layout");
    hip_check(hipStreamSynchronize(h->stream), "sync");
    return 0;
  } catch (const std::exception& e) {
    return fail(e, err, errlen);
  }
}

void dust2_system_discrete_free(dust2_system* h) {
  if (handle_ok(h, TAG_SYSTEM)) delete h;
}

// ============================================================================
// Test entry points
// ============================================================================
int dust2_amd_resample(const double* log_weights, size_t n, double u, uint32_t* idx,
                       double* out3, char* err, size_t errlen) {
  try {
    const size_t n_max = static_cast<size_t>(kMbMaxBlocks) * 8192;
    if (n < 1 || n > n_max) {
      throw std::runtime_error("n must lie in [1, " + std::to_string(n_max) + "]");
    }
    hipStream_t s = nullptr;
    dbuf dw(n * sizeof(double)), di(n * sizeof(uint32_t)), dl(3 * sizeof(double));
    to_device(dw, log_weights, n * sizeof(double), s);
    if (n <= 2048) {
      // one workgroup, the filter kernels' own routine
      hip_check(launch_resample_test(dw.get<double>(), static_cast<int>(n), u,
                                     di.get<uint32_t>(), dl.get<double>(), s),
                "resample kernel");
    } else {
      // many workgroups, partial results and cumulative sums in scratch
      dbuf scratch(mb_resample_scratch_doubles(static_cast<int>(n)) * sizeof(double));
      mb_resample task{};
      task.w = dw.get<double>();
      task.idx = di.get<uint32_t>();
      task.u = u;
      task.out3 = dl.get<double>();
      task.scratch = scratch.get<double>();
      task.P = static_cast<int>(n);
      hip_check(launch_mb_resample_one(task, 0, s), "resample kernels");
      hip_check(hipStreamSynchronize(s), "sync");
    }
    to_host(idx, di, n * sizeof(uint32_t), s);
    to_host(out3, dl, 3 * sizeof(double), s);
    return 0;
  } catch (const std::exception& e) {
    return fail(e, err, errlen);
  }
}

int dust2_amd_resample_noalloc(const double* d_log_weights, size_t n, double u, uint32_t* d_idx,
                               double* d_out3, void* stream, char* err, size_t errlen) {
  try {
    if (n < 1 || n > kNoAllocMaxParticles) {
      throw std::runtime_error("n must lie in [1, " + std::to_string(kNoAllocMaxParticles) +
                               "] (the allocation-free scratch holds " +
                               std::to_string(kNoAllocMaxParticles) + " weights)");
    }
    if (d_log_weights == nullptr || d_idx == nullptr || d_out3 == nullptr) {
      throw std::runtime_error("the weights, indices and results must be device buffers");
    }
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (n <= 2048) {
      hip_check(launch_resample_test(d_log_weights, static_cast<int>(n), u, d_idx, d_out3, s),
                "resample kernel");
    } else {
      double* scratch = noalloc_scratch_ptr();
      if (mb_resample_scratch_doubles(static_cast<int>(n)) > kNoAllocScratchDoubles) {
        throw std::runtime_error("allocation-free scratch too small");
      }
      mb_resample task{};
      task.w = d_log_weights;
      task.idx = d_idx;
      task.u = u;
      task.out3 = d_out3;
      task.scratch = scratch;
      task.P = static_cast<int>(n);
      hip_check(launch_mb_resample_one(task, 0, s), "resample kernels");
    }
    return 0;
  } catch (const std::exception& e) {
    return fail(e, err, errlen);
  }
}

int dust2_amd_trajectory_lineage(const double* raw, const uint32_t* ancestry, size_t n_time,
                                 size_t n_groups, size_t n_particles, size_t n_state,
                                 const size_t* select, double* out, char* err, size_t errlen) {
  try {
    if (n_time == 0 || n_groups == 0 || n_particles == 0 || n_state == 0) {
      throw std::runtime_error("sizes must be positive");
    }
    const size_t T = n_time, G = n_groups, P = n_particles, S = n_state;
    // caller layout [t][g][p][s] -> device [g][t][s][p]; ancestry [t][g][p] -> [g][t][p]
    std::vector<double> traj(T * G * P * S);
    std::vector<uint32_t> anc(ancestry ? T * G * P : 0);
    for (size_t t = 0; t < T; ++t) {
      for (size_t g = 0; g < G; ++g) {
        for (size_t p = 0; p < P; ++p) {
          for (size_t c = 0; c < S; ++c) {
            traj[((g * T + t) * S + c) * P + p] = raw[((t * G + g) * P + p) * S + c];
          }
          if (ancestry) {
            const uint32_t a = ancestry[(t * G + g) * P + p];
            if (a >= P) throw std::runtime_error("ancestor index out of range");
            anc[(g * T + t) * P + p] = a;
          }
        }
      }
    }
    std::vector<uint32_t> groups(G), sel(G);
    for (size_t g = 0; g < G; ++g) {
      groups[g] = static_cast<uint32_t>(g);
      if (select) {
        if (select[g] >= P) throw std::runtime_error("selected particle out of range");
        sel[g] = static_cast<uint32_t>(select[g]);
      }
    }
    hipStream_t s = nullptr;
    const size_t p_out = select ? 1 : P;
    dbuf dt(traj.size() * sizeof(double)), da(std::max<size_t>(anc.size(), 1) * 4),
        dg(G * 4), ds(G * 4), dout(T * G * p_out * S * sizeof(double));
    to_device(dt, traj.data(), traj.size() * sizeof(double), s);
    if (ancestry) to_device(da, anc.data(), anc.size() * 4, s);
    to_device(dg, groups.data(), G * 4, s);
    to_device(ds, sel.data(), G * 4, s);
    hip_check(launch_lineage(dt.get(), ancestry ? da.get<uint32_t>() : nullptr,
                             dg.get<uint32_t>(), static_cast<int>(G), static_cast<int>(T),
                             static_cast<int>(S), static_cast<int>(P),
                             select ? ds.get<uint32_t>() : nullptr, nullptr, nullptr, 0,
                             true, dout.get<double>(), s),
              "lineage kernel");
    to_host(out, dout, T * G * p_out * S * sizeof(double), s);
    return 0;
  } catch (const std::exception& e) {
    return fail(e, err, errlen);
  }
}

int dust2_amd_interpolate(int kind, const double* t, size_t n_t, const double* y,
                          size_t n_y, size_t n_series, const double* z, size_t n_z, double* out,
                          char* err, size_t errlen) {
  try {
    if (kind != INTERP_CONSTANT && kind != INTERP_LINEAR && kind != INTERP_SPLINE) {
      throw std::runtime_error("unknown interpolation kind");
    }
    if (n_series == 0 || n_y % n_series != 0) {
      throw std::runtime_error("'y' must hold n_series values per time");
    }
    interp_validate(t, n_t, n_y / n_series, "t", "y");
    if (n_t == 0 || (kind != INTERP_CONSTANT && n_t < 2)) {
      throw std::runtime_error("Interpolation needs at least " +
                               std::string(kind == INTERP_CONSTANT ? "one time" : "two times"));
    }
    const std::vector<double> k = kind == INTERP_SPLINE
                                      ? interp_spline_slopes(t, y, n_t, n_series)
                                      : std::vector<double>(n_y, 0.0);
    hipStream_t s = nullptr;
    dbuf dt(n_t * 8), dy(n_y * 8), dk(n_y * 8), dz(std::max<size_t>(n_z, 1) * 8),
        dout(std::max<size_t>(n_z * n_series, 1) * 8), derr(std::max<size_t>(n_z, 1) * 4);
    to_device(dt, t, n_t * 8, s);
    to_device(dy, y, n_y * 8, s);
    to_device(dk, k.data(), n_y * 8, s);
    if (n_z) to_device(dz, z, n_z * 8, s);
    hip_check(launch_interpolate(kind, dt.get<double>(), dy.get<double>(), dk.get<double>(),
                                 static_cast<int>(n_t), static_cast<int>(n_series),
                                 dz.get<double>(), static_cast<int>(n_z), dout.get<double>(),
                                 derr.get<int>(), s),
              "interpolate kernel");
    std::vector<int> e(n_z);
    if (n_z) {
      to_host(out, dout, n_z * n_series * 8, s);
      to_host(e.data(), derr, n_z * 4, s);
    }
    for (size_t q = 0; q < n_z; ++q) {
      if (e[q] == INTERP_OK) continue;
      // dust2/interpolate/interpolate.hpp:20-40, stream formatting
      std::ostringstream msg;
      if (e[q] == INTERP_BEFORE) {
        msg << "Tried to interpolate at time = " << z[q] << ", which is " << t[0] - z[q]
            << " before the first time (" << t[0] << ")";
      } else {
        msg << "Tried to interpolate at time = " << z[q] << ", which is "
            << z[q] - t[n_t - 1] << " after the last time (" << t[n_t - 1] << ")";
      }
      throw std::runtime_error(msg.str());
    }
    return 0;
  } catch (const std::exception& e) {
    return fail(e, err, errlen);
  }
}

int dust2_amd_rng_streams(const uint64_t* seed_state, size_t n_seed_streams, uint64_t seed,
                          size_t n_streams, uint64_t* out, char* err, size_t errlen) {
  try {
    hipStream_t s = nullptr;
    dbuf d(n_streams * 32);
    expand_streams(seed_state, n_seed_streams, seed, n_streams, d.get<uint64_t>(), s);
    to_host(out, d, n_streams * 32, s);
    return 0;
  } catch (const std::exception& e) {
    return fail(e, err, errlen);
  }
}

int dust2_amd_valu_probe(int precision, int waves_per_simd, int iters, double* ms_out,
                         double* rate_out, int* n_simd_out, char* err, size_t errlen) {
  try {
    if (waves_per_simd < 1 || waves_per_simd > 8) {
      throw std::runtime_error("'waves_per_simd' must be between 1 and 8");
    }
    if (iters < 1) throw std::runtime_error("'iters' must be at least 1");
    const bool fp64 = precision == DUST2_FP64;
    int dev = 0, n_cu = 0;
    hip_check(hipGetDevice(&dev), "hipGetDevice");
    hip_check(hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev),
              "hipDeviceGetAttribute");
    // four wavefronts per workgroup, one per SIMD of a compute unit
    const int n_blocks = n_cu * waves_per_simd;
    hipStream_t s = nullptr;
    hip_check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate");
    dbuf out(static_cast<size_t>(n_blocks) * 256 * sizeof(double));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hip_check(hipEventCreate(&e0), "hipEventCreate");
    hip_check(hipEventCreate(&e1), "hipEventCreate");
    hip_check(launch_valu_probe(fp64, n_blocks, 10, out.get<double>(), s), "probe kernel");
    hip_check(hipEventRecord(e0, s), "hipEventRecord");
    hip_check(launch_valu_probe(fp64, n_blocks, iters, out.get<double>(), s), "probe kernel");
    hip_check(hipEventRecord(e1, s), "hipEventRecord");
    hip_check(hipEventSynchronize(e1), "hipEventSynchronize");
    float ms = 0;
    hip_check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipStreamDestroy(s);
    const double wave_instr = 4.0 * n_blocks * static_cast<double>(iters) *
                              valu_probe_fmas_per_iteration(fp64);
    *ms_out = ms;
    *rate_out = wave_instr / (4.0 * n_cu) / (ms * 1e6);
    *n_simd_out = 4 * n_cu;
    return 0;
  } catch (const std::exception& e) {
    return fail(e, err, errlen);
  }
}

int dust2_amd_sample(int kind, int precision, uint64_t* rng, size_t n_streams, size_t n_draws,
                     const double* a, const double* b, int deterministic, double* out,
                     uint32_t* err_out, char* err, size_t errlen) {
  try {
    if (kind < 0 || kind > 4) throw std::runtime_error("unknown sampler kind");
    hipStream_t s = nullptr;
    dbuf dr(n_streams * 32), da(n_streams * 8), db(n_streams * 8),
        dout(n_streams * n_draws * 8), derr(n_streams * 4);
    to_device(dr, rng, n_streams * 32, s);
    to_device(da, a, n_streams * 8, s);
    to_device(db, b, n_streams * 8, s);
    hip_check(launch_sample_test(kind, precision == DUST2_FP64, dr.get<uint64_t>(),
                                 static_cast<int>(n_streams), static_cast<int>(n_draws),
                                 da.get<double>(), db.get<double>(), deterministic,
                                 dout.get<double>(), derr.get<uint32_t>(), s),
              "sample kernel");
    to_host(out, dout, n_streams * n_draws * 8, s);
    to_host(err_out, derr, n_streams * 4, s);
    to_host(rng, dr, n_streams * 32, s);
    return 0;
  } catch (const std::exception& e) {
    return fail(e, err, errlen);
  }
}

}  // extern "C"

// ============================================================================
// Unfilter (deterministic likelihood, SIR ODE, Dormand-Prince 5(4))
// ============================================================================
struct dust2_unfilter {
  uint32_t tag = TAG_UNFILTER;
  bool fp64 = true;
  size_t G = 0, P = 0, n_time = 0, n_groups_data = 0;
  double time_start = 0;
  std::vector<double> pars;  // G x 4
  ode_control_dev control{};
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double last_ms = 0;
  dbuf d_critical;
  dbuf d_pars, d_times, d_obs, d_lfact, d_state, d_stats, d_errors, d_counters, d_active, d_grad,
      d_traj;
  // the error count and the log-likelihoods of a run come back in one
  // stream synchronisation through pinned memory
  pinned h_result;
  std::vector<size_t> active_idx;     // groups whose flags d_active holds
  std::vector<double> times;          // data times
  saved_run saved;                    // trajectories and snapshots of the last run
  std::vector<uint8_t> grad_current;  // per group: the gradient of the last run is valid
  std::vector<size_t> last_index;     // groups of the last run
  size_t real_size() const { return fp64 ? 8 : 4; }
  ~dust2_unfilter() {
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace {
constexpr size_t kSirodePars = 4;

// dust_ode_control() values (R/interface-continuous.R:51-79: max_steps,
// atol, rtol, step_size_min, step_size_max; null for the defaults) with
// the C++ controller constants (dust2/continuous/control.hpp:11-24).
ode_control_dev ode_control_from(const double* control) {
  const double def[5] = {10000, 1e-6, 1e-6, 0, std::numeric_limits<double>::infinity()};
  const double* c = control ? control : def;
  if (!(c[0] >= 1) || !(c[1] > 0) || !(c[2] > 0) || !(c[3] >= 0) || !(c[4] >= 0)) {
    throw std::runtime_error("Invalid ODE control");
  }
  ode_control_dev ctl{};
  ctl.max_steps = static_cast<uint32_t>(std::min(c[0], 4e9));
  ctl.atol = c[1];
  ctl.rtol = c[2];
  ctl.step_size_min = std::max(c[3], std::numeric_limits<double>::epsilon());
  ctl.step_size_max = c[4];
  ctl.factor_safe = 0.9;
  ctl.factor_min = 0.2;
  ctl.factor_max = 10.0;
  ctl.beta = 0.04;
  ctl.constant = 0.2 - 0.04 * 0.75;
  return ctl;
}

// control.critical_times (dust2/continuous/control.hpp:26-39): the times
// must increase; kept on the device for the kernels.
void set_critical_times(ode_control_dev& ctl, dbuf& d, const double* times, size_t n,
                        hipStream_t s) {
  for (size_t i = 1; i < n; ++i) {
    if (!(times[i] > times[i - 1])) {
      throw std::runtime_error("'critical_times' must be strictly increasing");
    }
  }
  d.alloc(std::max<size_t>(n, 1) * sizeof(double));
  if (n) to_device(d, times, n * sizeof(double), s);
  hip_check(hipStreamSynchronize(s), "sync");
  ctl.critical_times = n ? d.get<double>() : nullptr;
  ctl.n_critical = static_cast<int>(n);
}

void check_sirode_pars(const double* row) {
  const char* names[] = {"N", "I0", "beta", "gamma"};
  for (size_t i = 0; i < kSirodePars; ++i) {
    if (!std::isfinite(row[i])) {
      throw std::runtime_error(std::string("Parameter '") + names[i] + "' must be finite");
    }
  }
}

void check_unfilter(dust2_unfilter* h) { check_handle(h, TAG_UNFILTER, "unfilter"); }
}  // namespace

extern "C" {

int dust2_unfilter_sirode_alloc(const double* pars, size_t n_groups, size_t n_particles,
                                double time_start, const double* time, size_t n_time,
                                const double* data, size_t n_groups_data, const double* control,
                                int precision, dust2_unfilter** out, char* err, size_t errlen) {
  try {
    check_sizes(n_groups, n_particles);
    check_time_sequence(time_start, time, n_time);
    if (n_groups_data != 1 && n_groups_data != n_groups) {
      throw std::runtime_error("Expected 'data' to have 1 or " + std::to_string(n_groups) +
                               " groups, but it had " + std::to_string(n_groups_data));
    }
    if (precision != DUST2_FP64 && precision != DUST2_FP32) {
      throw std::runtime_error("'precision' must be DUST2_FP64 or DUST2_FP32");
    }
    if (n_groups * n_particles > (size_t(1) << 31)) {
      throw std::runtime_error("too many particles");
    }
    for (size_t g = 0; g < n_groups; ++g) check_sirode_pars(pars + g * kSirodePars);
    auto h = std::make_unique<dust2_unfilter>();
    h->fp64 = precision == DUST2_FP64;
    h->G = n_groups;
    h->P = n_particles;
    h->n_time = n_time;
    h->n_groups_data = n_groups_data;
    h->time_start = time_start;
    h->times.assign(time, time + n_time);
    h->pars.assign(pars, pars + n_groups * kSirodePars);
    h->grad_current.assign(n_groups, 0);
    h->control = ode_control_from(control);
    hip_check(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking), "hipStreamCreate");
    hip_check(hipEventCreate(&h->ev0), "hipEventCreate");
    hip_check(hipEventCreate(&h->ev1), "hipEventCreate");
    const size_t n_total = n_groups * n_particles;
    const size_t nd = n_time * n_groups_data;
    std::vector<double> lfact(nd);
    for (size_t i = 0; i < nd; ++i) lfact[i] = std::isnan(data[i]) ? 0.0 : std::lgamma(data[i] + 1);
    h->d_pars.alloc(h->pars.size() * sizeof(double));
    h->d_times.alloc(n_time * sizeof(double));
    h->d_obs.alloc(nd * sizeof(double));
    h->d_lfact.alloc(nd * sizeof(double));
    h->d_state.alloc(n_total * 5 * h->real_size());
    h->d_stats.alloc(n_total * 4 * sizeof(double));
    h->d_errors.alloc(n_total * sizeof(error_record));
    h->d_counters.alloc(4 * sizeof(uint32_t));
    h->d_active.alloc(n_groups);
    to_device(h->d_pars, h->pars.data(), h->pars.size() * sizeof(double), h->stream);
    to_device(h->d_times, time, n_time * sizeof(double), h->stream);
    to_device(h->d_obs, data, nd * sizeof(double), h->stream);
    to_device(h->d_lfact, lfact.data(), nd * sizeof(double), h->stream);
    hip_check(hipMemsetAsync(h->d_errors.get(), 0, h->d_errors.bytes(), h->stream), "memset");
    hip_check(hipMemsetAsync(h->d_state.get(), 0, h->d_state.bytes(), h->stream), "memset");
    hip_check(hipStreamSynchronize(h->stream), "sync");
    *out = h.release();
    return 0;
  } catch (const std::exception& e) {
    return fail(e, err, errlen);
  }
}

}  // extern "C"

namespace {
// One launch of the DP5 unfilter over the groups in idx; sens_param >= 0
// runs the forward sensitivity of that parameter into d_grad as well.
void sirode_launch(dust2_unfilter* h, const std::vector<size_t>& idx, int sens_param,
                   bool save_traj) {
  hipStream_t s = h->stream;
  // every group: no flags; otherwise the flags, uploaded when they change
  const bool all = idx.size() == h->G;
  if (!all && idx != h->active_idx) {
    std::vector<uint8_t> active(h->G, 0);
    for (size_t g : idx) active[g] = 1;
    to_device(h->d_active, active.data(), h->G, s);
    h->active_idx = idx;
  }
  ode_launch a{};
  a.sens_param = sens_param;
  a.grad = sens_param >= 0 ? h->d_grad.get<double>() : nullptr;
  if (save_traj && h->d_traj.bytes() == 0) {
    h->d_traj.alloc(h->n_time * 5 * h->G * h->P * h->real_size());
  }
  a.traj = save_traj ? h->d_traj.get() : nullptr;
  a.pars = h->d_pars.get<double>();
  a.active = all ? nullptr : h->d_active.get<uint8_t>();
  a.n_particles = static_cast<int>(h->P);
  a.n_total = static_cast<int>(h->G * h->P);
  a.time_start = h->time_start;
  a.times = h->d_times.get<double>();
  a.n_time = static_cast<int>(h->n_time);
  a.obs = h->d_obs.get<double>();
  a.lfact = h->d_lfact.get<double>();
  a.obs_stride = h->n_groups_data;
  a.obs_group_stride = h->n_groups_data == 1 ? 0 : 1;
  a.control = h->control;
  a.state = h->d_state.get();
  // each particle's log-likelihood is written once, at the end of its
  // run, straight into pinned host memory (no copy after the kernel)
  const size_t n_ll = h->G * h->P;
  if (h->h_result.bytes() < 256 + n_ll * sizeof(double)) {
    h->h_result.alloc(256 + n_ll * sizeof(double));
  }
  a.ll = reinterpret_cast<double*>(h->h_result.get<char>() + 256);
  a.stats = h->d_stats.get<double>();
  a.errors = h->d_errors.get<error_record>();
  // the error flag also lives in pinned memory: cleared here by the host
  // before the launch is queued, set by a failing particle, read after the
  // synchronisation (no memset or copy around the kernel)
  volatile uint32_t* flag = h->h_result.get<uint32_t>();
  *flag = 0;
  a.error_flag = h->h_result.get<uint32_t>();
  hip_check(hipEventRecord(h->ev0, s), "hipEventRecord");
  hip_check(launch_dp5_unfilter(a, h->fp64, s), "dp5 unfilter kernel");
  hip_check(hipEventRecord(h->ev1, s), "hipEventRecord");
  hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
  float ms = 0;
  hip_check(hipEventElapsedTime(&ms, h->ev0, h->ev1), "hipEventElapsedTime");
  h->last_ms = ms;
  if (*flag != 0) {
    std::vector<error_record> rec(h->G * h->P);
    to_host(rec.data(), h->d_errors, rec.size() * sizeof(error_record), s);
    size_t count = 0;
    for (const error_record& r : rec) count += r.code != ERR_NONE ? 1 : 0;
    hip_check(hipMemsetAsync(h->d_errors.get(), 0, h->d_errors.bytes(), s), "memset");
    hip_check(hipStreamSynchronize(s), "sync");
    throw std::runtime_error(error_report(rec, count));
  }
}

// The log-likelihoods of the last launch, already in pinned memory.
void sirode_copy_ll(dust2_unfilter* h, const std::vector<size_t>& idx, double* ll_out) {
  const double* ll = reinterpret_cast<const double*>(h->h_result.get<char>() + 256);
  if (idx.size() == h->G) {
    std::memcpy(ll_out, ll, h->G * h->P * sizeof(double));
    return;
  }
  for (size_t i = 0; i < idx.size(); ++i) {
    std::memcpy(ll_out + i * h->P, ll + idx[i] * h->P, h->P * sizeof(double));
  }
}
}  // namespace

extern "C" {

int dust2_unfilter_sirode_run(dust2_unfilter* h, int save_trajectories,
                              const double* save_snapshots, size_t n_snapshots,
                              const size_t* index_state, size_t n_index_state,
                              const size_t* index_group, size_t n_index, double* ll_out,
                              char* err, size_t errlen) {
  try {
    check_unfilter(h);
    const auto idx = check_index(index_group, n_index, h->G);
    saved_run req = check_save(save_trajectories, save_snapshots, n_snapshots, index_state,
                               n_index_state, 5, h->times);
    for (size_t g : idx) h->grad_current[g] = 0;
    h->saved = saved_run{};
    sirode_launch(h, idx, -1, req.any());
    h->saved = std::move(req);
    sirode_copy_ll(h, idx, ll_out);
    h->last_index = idx;
    return 0;
  } catch (const std::exception& e) {
    return fail(e, err, errlen);
  }
}

int dust2_unfilter_sirode_run_gradient(dust2_unfilter* h, int save_trajectories,
                                       const double* save_snapshots, size_t n_snapshots,
                                       const size_t* index_state, size_t n_index_state,
                                       const size_t* index_group, size_t n_index,
                                       double* ll_out, char* err, size_t errlen) {
  try {
    check_unfilter(h);
    if (!h->fp64) {
      throw std::runtime_error("The ODE gradient is computed in FP64 only");
    }
    const auto idx = check_index(index_group, n_index, h->G);
    saved_run req = check_save(save_trajectories, save_snapshots, n_snapshots, index_state,
                               n_index_state, 5, h->times);
    for (size_t g : idx) h->grad_current[g] = 0;
    if (h->d_grad.bytes() == 0) h->d_grad.alloc(h->G * h->P * 3 * sizeof(double));
    double total_ms = 0;
    h->saved = saved_run{};
    for (int param = 0; param < 3; ++param) {
      sirode_launch(h, idx, param, param == 0 && req.any());
      total_ms += h->last_ms;
    }
    h->saved = std::move(req);
    h->last_ms = total_ms;
    sirode_copy_ll(h, idx, ll_out);
    for (size_t g : idx) h->grad_current[g] = 1;
    h->last_index = idx;
    return 0;
  } catch (const std::exception& e) {
    return fail(e, err, errlen);
  }
}

int dust2_unfilter_sirode_last_gradient(dust2_unfilter* h, double* out, char* err,
                                        size_t errlen) {
  try {
    check_unfilter(h);
    bool any = false;
    for (size_t g : h->last_index) any = any || h->grad_current[g];
    if (!any) {
      throw std::runtime_error("System was not run with 'adjoint = TRUE'");
    }
    std::vector<double> grad(h->G * h->P * 3);
    to_host(grad.data(), h->d_grad, grad.size() * sizeof(double), h->stream);
    for (size_t i = 0; i < h->last_index.size(); ++i) {
      std::memcpy(out + 3 * h->P * i, grad.data() + 3 * h->P * h->last_index[i],
                  3 * h->P * sizeof(double));
    }
    return 0;
  } catch (const std::exception& e) {
    return fail(e, err, errlen);
  }
}

}  // extern "C"

namespace {
// trajectories (snapshots = false) or snapshots of the last run of the ODE
// unfilter: device [t][S][G*P] -> out [t or snapshot][groups][P][states]
void sirode_export(dust2_unfilter* h, bool snapshots, double* out) {
  const size_t NT = h->G * h->P, T = h->n_time;
  std::vector<char> tmp(T * 5 * NT * h->real_size());
  to_host(tmp.data(), h->d_traj, tmp.size(), h->stream);
  const std::vector<int32_t>* slot = snapshots ? &h->saved.slot : nullptr;
  const std::vector<uint32_t> all;
  const std::vector<uint32_t>& states = snapshots ? all : h->saved.states;
  if (h->fp64) {
    export_saved_states(reinterpret_cast<const double*>(tmp.data()), T, 5, NT, h->P,
                        h->last_index, slot, states, out);
  } else {
    export_saved_states(reinterpret_cast<const float*>(tmp.data()), T, 5, NT, h->P,
                        h->last_index, slot, states, out);
  }
}
}  // namespace

extern "C" {

int dust2_unfilter_sirode_last_trajectories(dust2_unfilter* h, double* out, char* err,
                                            size_t errlen) {
  try {
    check_unfilter(h);
    if (!h->saved.traj) throw std::runtime_error("Trajectories are not current");
    sirode_export(h, false, out);
    return 0;
  } catch (const std::exception& e) {
    return fail(e, err, errlen);
  }
}

int dust2_unfilter_sirode_last_snapshots(dust2_unfilter* h, double* out, char* err,
                                         size_t errlen) {
  try {
    check_unfilter(h);
    if (h->saved.n_snap == 0) throw std::runtime_error("Snapshots are not current");
    sirode_export(h, true, out);
    return 0;
  } catch (const std::exception& e) {
    return fail(e, err, errlen);
  }
}

int dust2_unfilter_sirode_update_pars(dust2_unfilter* h, const double* pars,
                                      const size_t* index_group, size_t n_index, char* err,
                                      size_t errlen) {
  try {
    check_unfilter(h);
    const auto idx = check_index(index_group, n_index, h->G);
    for (size_t i = 0; i < idx.size(); ++i) {
      const double* row = pars + i * kSirodePars;
      check_sirode_pars(row);
      double* dst = &h->pars[idx[i] * kSirodePars];
      dst[1] = row[1];
      dst[2] = row[2];
      dst[3] = row[3];
      h->grad_current[idx[i]] = 0;
    }
    to_device(h->d_pars, h->pars.data(), h->pars.size() * sizeof(double), h->stream);
    hip_check(hipStreamSynchronize(h->stream), "sync");
    return 0;
  } catch (const std::exception& e) {
    return fail(e, err, errlen);
  }
}

int dust2_unfilter_sirode_last_state(dust2_unfilter* h, double* out, char* err, size_t errlen) {
  try {
    check_unfilter(h);
    const size_t n = h->G * h->P;
    std::vector<char> tmp(h->d_state.bytes());
    to_host(tmp.data(), h->d_state, tmp.size(), h->stream);
    for (size_t k = 0; k < n; ++k) {
      for (size_t c = 0; c < 5; ++c) {
        out[k * 5 + c] = h->fp64 ? reinterpret_cast<const double*>(tmp.data())[c * n + k]
                                 : reinterpret_cast<const float*>(tmp.data())[c * n + k];
      }
    }
    return 0;
  } catch (const std::exception& e) {
    return fail(e, err, errlen);
  }
}

int dust2_unfilter_sirode_last_stats(dust2_unfilter* h, double* out, double* kernel_ms,
                                     char* err, size_t errlen) {
  try {
    check_unfilter(h);
    to_host(out, h->d_stats, h->G * h->P * 4 * sizeof(double), h->stream);
    if (kernel_ms) *kernel_ms = h->last_ms;
    return 0;
  } catch (const std::exception& e) {
    return fail(e, err, errlen);
  }
}

void dust2_unfilter_sirode_free(dust2_unfilter* h) {
  if (handle_ok(h, TAG_UNFILTER)) delete h;
}

int dust2_unfilter_sirode_set_critical_times(dust2_unfilter* h, const double* times, size_t n,
                                             char* err, size_t errlen) {
  try {
    check_unfilter(h);
    set_critical_times(h->control, h->d_critical, times, n, h->stream);
    return 0;
  } catch (const std::exception& e) {
    return fail(e, err, errlen);
  }
}

}  // extern "C"

// ============================================================================
// Continuous-time system: the SIR ODE (dust2/continuous/system.hpp)
// ============================================================================
struct dust2_ode_system {
  uint32_t tag = TAG_ODE_SYSTEM;
  size_t G = 0, P = 0;
  double time = 0;
  std::vector<double> pars;           // G x 4
  std::vector<uint8_t> reinit;        // per group: the solver must initialise
  ode_control_dev control{};
  uint32_t record_cap = 0;            // step times kept per particle (0: off)
  hipStream_t stream = nullptr;
  dbuf d_pars, d_state, d_internals, d_reinit, d_targets, d_out, d_errors, d_counters,
      d_critical, d_step_times, d_step_count, d_tmp;
  bool errors_pending = false;
  ~dust2_ode_system() {
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace {

void check_ode_system(dust2_ode_system* h, const char* action) {
  check_handle(h, TAG_ODE_SYSTEM, "system");
  if (h->errors_pending) {
    throw std::runtime_error(std::string("Can't currently ") + action +
                             " this system: errors are pending");
  }
}

// Integrates every particle from h->time through `targets`, writing the
// state after each target to out ([n][G][P][S]) when given.
void ode_system_run(dust2_ode_system* h, const double* targets, size_t n, double* out) {
  hipStream_t s = h->stream;
  const size_t n_total = h->G * h->P;
  h->d_targets.alloc(std::max<size_t>(n, 1) * sizeof(double));
  if (n) to_device(h->d_targets, targets, n * sizeof(double), s);
  to_device(h->d_reinit, h->reinit.data(), h->G, s);
  if (out) h->d_out.alloc(std::max<size_t>(n, 1) * kSirState * n_total * sizeof(double));
  hip_check(hipMemsetAsync(h->d_counters.get(), 0, h->d_counters.bytes(), s), "memset");
  ode_sys_launch a{};
  a.pars = h->d_pars.get<double>();
  a.n_particles = static_cast<int>(h->P);
  a.n_total = static_cast<int>(n_total);
  a.reinit = h->d_reinit.get<uint8_t>();
  a.time = h->time;
  a.targets = h->d_targets.get<double>();
  a.n_targets = static_cast<int>(n);
  a.control = h->control;
  a.state = h->d_state.get<double>();
  a.internals = h->d_internals.get<ode_internals>();
  a.out = out ? h->d_out.get<double>() : nullptr;
  a.record = ode_record{h->record_cap ? h->d_step_times.get<double>() : nullptr,
                        h->d_step_count.get<uint32_t>(), h->record_cap};
  a.errors = h->d_errors.get<error_record>();
  a.error_count = h->d_counters.get<uint32_t>();
  hip_check(launch_dp5_system(a, s), "ode system kernel");
  uint32_t count = 0;
  to_host(&count, h->d_counters, sizeof(count), s);
  std::fill(h->reinit.begin(), h->reinit.end(), 0);
  if (count != 0) {
    std::vector<error_record> rec(n_total);
    to_host(rec.data(), h->d_errors, rec.size() * sizeof(error_record), s);
    hip_check(hipMemsetAsync(h->d_errors.get(), 0, h->d_errors.bytes(), s), "memset");
    hip_check(hipStreamSynchronize(s), "sync");
    h->errors_pending = true;
    throw std::runtime_error(error_report(rec, count));
  }
  if (n) h->time = targets[n - 1];
  if (out) {
    const size_t block = kSirState * n_total;
    std::vector<double> tmp(n * block);
    to_host(tmp.data(), h->d_out, tmp.size() * sizeof(double), s);
    for (size_t q = 0; q < n; ++q) {
      soa_to_aos(tmp.data() + q * block, out + q * block, 1, n_total, kSirState, true);
    }
  }
}

}  // namespace

extern "C" {

int dust2_system_sirode_alloc(const double* pars, size_t n_groups, size_t n_particles,
                              double time, const double* control, const double* critical_times,
                              size_t n_critical, size_t record_step_times,
                              dust2_ode_system** out, char* err, size_t errlen) {
  try {
    check_sizes(n_groups, n_particles);
    for (size_t g = 0; g < n_groups; ++g) check_sirode_pars(pars + g * kSirodePars);
    auto h = std::make_unique<dust2_ode_system>();
    h->G = n_groups;
    h->P = n_particles;
    h->time = time;
    h->pars.assign(pars, pars + n_groups * kSirodePars);
    h->reinit.assign(n_groups, 1);
    h->control = ode_control_from(control);
    h->record_cap = static_cast<uint32_t>(std::min<size_t>(record_step_times, 1u << 20));
    hip_check(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking), "hipStreamCreate");
    const size_t n_total = n_groups * n_particles;
    h->d_pars.alloc(n_groups * kSirodePars * sizeof(double));
    to_device(h->d_pars, pars, n_groups * kSirodePars * sizeof(double), h->stream);
    h->d_state.alloc(kSirState * n_total * sizeof(double));
    h->d_internals.alloc(n_total * sizeof(ode_internals));
    h->d_reinit.alloc(n_groups);
    h->d_errors.alloc(n_total * sizeof(error_record));
    h->d_counters.alloc(4 * sizeof(uint32_t));
    h->d_step_count.alloc(n_total * sizeof(uint32_t));
    if (h->record_cap) h->d_step_times.alloc(n_total * h->record_cap * sizeof(double));
    hip_check(hipMemsetAsync(h->d_state.get(), 0, h->d_state.bytes(), h->stream), "memset");
    hip_check(hipMemsetAsync(h->d_internals.get(), 0, h->d_internals.bytes(), h->stream),
              "memset");
    hip_check(hipMemsetAsync(h->d_errors.get(), 0, h->d_errors.bytes(), h->stream), "memset");
    hip_check(hipMemsetAsync(h->d_step_count.get(), 0, h->d_step_count.bytes(), h->stream),
              "memset");
    set_critical_times(h->control, h->d_critical, critical_times, n_critical, h->stream);
    *out = h.release();
    return 0;
  } catch (const std::exception& e) {
    return fail(e, err, errlen);
  }
}

int dust2_system_sirode_run_to_time(dust2_ode_system* h, double time, char* err,
                                    size_t errlen) {
  try {
    check_ode_system(h, "run");
    if (time < h->time) {
      char buf[256];
      std::snprintf(buf, sizeof buf, "Can't run to time %f, system already at time %f", time,
                    h->time);
      throw std::runtime_error(buf);
    }
    ode_system_run(h, &time, 1, nullptr);
    return 0;
  } catch (const std::exception& e) {
    return fail(e, err, errlen);
  }
}

int dust2_system_sirode_simulate(dust2_ode_system* h, const double* times, size_t n_times,
                                 double* out, char* err, size_t errlen) {
  try {
    check_ode_system(h, "simulate");
    double prev = h->time;
    for (size_t i = 0; i < n_times; ++i) {
      if (times[i] < prev) {
        throw std::runtime_error("Expected 'times' to be increasing, starting at the "
                                 "system's time");
      }
      prev = times[i];
    }
    ode_system_run(h, times, n_times, out);
    return 0;
  } catch (const std::exception& e) {
    return fail(e, err, errlen);
  }
}

int dust2_system_sirode_state(dust2_ode_system* h, double* out, char* err, size_t errlen) {
  try {
    check_ode_system(h, "get state from");
    std::vector<double> tmp(kSirState * h->G * h->P);
    to_host(tmp.data(), h->d_state, tmp.size() * sizeof(double), h->stream);
    soa_to_aos(tmp.data(), out, 1, h->G * h->P, kSirState, true);
    return 0;
  } catch (const std::exception& e) {
    return fail(e, err, errlen);
  }
}

int dust2_system_sirode_set_state(dust2_ode_system* h, const double* state, char* err,
                                  size_t errlen) {
  try {
    check_handle(h, TAG_ODE_SYSTEM, "system");  // allowed while errors are pending
    std::vector<double> tmp(kSirState * h->G * h->P);
    aos_to_soa(state, tmp.data(), 1, h->G * h->P, kSirState, true);
    to_device(h->d_state, tmp.data(), tmp.size() * sizeof(double), h->stream);
    hip_check(hipStreamSynchronize(h->stream), "sync");
    std::fill(h->reinit.begin(), h->reinit.end(), 1);
    h->errors_pending = false;
    return 0;
  } catch (const std::exception& e) {
    return fail(e, err, errlen);
  }
}

int dust2_system_sirode_set_state_initial(dust2_ode_system* h, char* err, size_t errlen) {
  try {
    check_handle(h, TAG_ODE_SYSTEM, "system");  // allowed while errors are pending
    // sirode.cpp:35-45: S = N - I0, I = I0, the rest 0
    std::vector<double> st(h->G * h->P * kSirState, 0.0);
    for (size_t g = 0; g < h->G; ++g) {
      for (size_t p = 0; p < h->P; ++p) {
        double* y = st.data() + (g * h->P + p) * kSirState;
        y[0] = h->pars[g * kSirodePars + 0] - h->pars[g * kSirodePars + 1];
        y[1] = h->pars[g * kSirodePars + 1];
      }
    }
    h->errors_pending = false;
    return dust2_system_sirode_set_state(h, st.data(), err, errlen);
  } catch (const std::exception& e) {
    return fail(e, err, errlen);
  }
}

int dust2_system_sirode_time(dust2_ode_system* h, double* out, char* err, size_t errlen) {
  try {
    check_ode_system(h, "get time from");
    *out = h->time;
    return 0;
  } catch (const std::exception& e) {
    return fail(e, err, errlen);
  }
}

int dust2_system_sirode_set_time(dust2_ode_system* h, double time, char* err, size_t errlen) {
  try {
    check_ode_system(h, "set time for");
    h->time = time;
    std::fill(h->reinit.begin(), h->reinit.end(), 1);
    return 0;
  } catch (const std::exception& e) {
    return fail(e, err, errlen);
  }
}

int dust2_system_sirode_update_pars(dust2_ode_system* h, const double* pars, char* err,
                                    size_t errlen) {
  try {
    check_handle(h, TAG_ODE_SYSTEM, "system");  // allowed while errors are pending
    for (size_t g = 0; g < h->G; ++g) check_sirode_pars(pars + g * kSirodePars);
    h->pars.assign(pars, pars + h->G * kSirodePars);
    to_device(h->d_pars, pars, h->pars.size() * sizeof(double), h->stream);
    hip_check(hipStreamSynchronize(h->stream), "sync");
    // update_shared marks the group's solver for initialisation
    // (dust2/continuous/system.hpp:256-265)
    std::fill(h->reinit.begin(), h->reinit.end(), 1);
    return 0;
  } catch (const std::exception& e) {
    return fail(e, err, errlen);
  }
}

int dust2_system_sirode_compare_data(dust2_ode_system* h, const double* data,
                                     size_t n_groups_data, double* out, char* err,
                                     size_t errlen) {
  try {
    check_ode_system(h, "compare data for");
    if (n_groups_data != 1 && n_groups_data != h->G) {
      throw std::runtime_error("Expected data for one group or every group");
    }
    std::vector<double> st(h->G * h->P * kSirState);
    char buf[512];
    if (dust2_system_sirode_state(h, st.data(), buf, sizeof buf) != 0) {
      throw std::runtime_error(buf);
    }
    // sirode.cpp:89-101: log Poisson(incidence; cases_inc), no draw
    for (size_t g = 0; g < h->G; ++g) {
      const double x = data[n_groups_data == 1 ? 0 : g];
      for (size_t p = 0; p < h->P; ++p) {
        const double lambda = st[(g * h->P + p) * kSirState + 4];
        double v = 0;
        if (!std::isnan(x)) {
          v = (x == 0 && lambda == 0) ? 0
              : lambda == 0           ? -std::numeric_limits<double>::infinity()
                                      : x * std::log(lambda) - lambda - std::lgamma(x + 1);
        }
        out[g * h->P + p] = v;
      }
    }
    return 0;
  } catch (const std::exception& e) {
    return fail(e, err, errlen);
  }
}

int dust2_system_sirode_reorder(dust2_ode_system* h, const uint32_t* index, char* err,
                                size_t errlen) {
  try {
    check_ode_system(h, "reorder");
    const size_t n_total = h->G * h->P;
    for (size_t k = 0; k < n_total; ++k) {
      if (index[k] >= h->P) throw std::runtime_error("reorder index out of range");
    }
    // the state and the solver internals move together
    // (dust2/continuous/system.hpp:172-199)
    std::vector<double> st(kSirState * n_total), st2(kSirState * n_total);
    std::vector<ode_internals> in(n_total), in2(n_total);
    to_host(st.data(), h->d_state, st.size() * sizeof(double), h->stream);
    to_host(in.data(), h->d_internals, in.size() * sizeof(ode_internals), h->stream);
    for (size_t g = 0; g < h->G; ++g) {
      for (size_t p = 0; p < h->P; ++p) {
        const size_t to = g * h->P + p, from = g * h->P + index[to];
        for (size_t c = 0; c < kSirState; ++c) st2[c * n_total + to] = st[c * n_total + from];
        in2[to] = in[from];
      }
    }
    to_device(h->d_state, st2.data(), st2.size() * sizeof(double), h->stream);
    to_device(h->d_internals, in2.data(), in2.size() * sizeof(ode_internals), h->stream);
    hip_check(hipStreamSynchronize(h->stream), "sync");
    return 0;
  } catch (const std::exception& e) {
    return fail(e, err, errlen);
  }
}

int dust2_system_sirode_internals(dust2_ode_system* h, double* stats, double* step_times,
                                  size_t* n_step_times, char* err, size_t errlen) {
  try {
    check_ode_system(h, "get internals from");
    const size_t n_total = h->G * h->P;
    std::vector<ode_internals> in(n_total);
    to_host(in.data(), h->d_internals, in.size() * sizeof(ode_internals), h->stream);
    for (size_t k = 0; k < n_total; ++k) {
      stats[4 * k + 0] = in[k].n_steps;
      stats[4 * k + 1] = in[k].n_accepted;
      stats[4 * k + 2] = in[k].n_rejected;
      stats[4 * k + 3] = in[k].step_size;
    }
    if (n_step_times) {
      std::vector<uint32_t> cnt(n_total);
      to_host(cnt.data(), h->d_step_count, cnt.size() * sizeof(uint32_t), h->stream);
      for (size_t k = 0; k < n_total; ++k) n_step_times[k] = h->record_cap ? cnt[k] : 0;
      if (step_times && h->record_cap) {
        to_host(step_times, h->d_step_times, n_total * h->record_cap * sizeof(double),
                h->stream);
      }
    }
    return 0;
  } catch (const std::exception& e) {
    return fail(e, err, errlen);
  }
}

void dust2_system_sirode_free(dust2_ode_system* h) {
  if (handle_ok(h, TAG_ODE_SYSTEM)) delete h;
}

}  // extern "C"

// ============================================================================
// General continuous-time systems: events, carried variables, delays
// (dust2/continuous/system.hpp over the models of ode_models.hpp)
// ============================================================================
struct dust2_ode_general {
  uint32_t tag = TAG_ODE_GENERAL;
  int model = 0;
  odem::model_info info{};
  size_t G = 0, P = 0, S = 0;          // S = n_var + n_special + n_output
  int words = 0, hist_words = 0;
  double time = 0;
  std::vector<double> pars;            // G x n_pars
  std::vector<uint8_t> reinit;
  ode_control_dev control{};
  uint32_t record_cap = 0, event_cap = 0, history_cap = 0;
  bool save_history = false;
  double dt = 0;                       // mixed time: update interval (0: none)
  bool deterministic = false;
  hipStream_t stream = nullptr;
  dbuf d_pars, d_state, d_internals, d_reinit, d_targets, d_out, d_errors, d_counters,
      d_critical, d_step_times, d_step_count, d_events, d_history, d_rng, d_data, d_ll,
      d_active;
  bool errors_pending = false;
  ~dust2_ode_general() {
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace {

const char* const kGeneralModels[odem::MODEL_COUNT] = {
    "bounce", "change", "eventzero", "specialevent", "sirdelay", "logdelay", "special",
    "malaria"};

int general_model_id(const char* name) {
  for (int i = 0; i < odem::MODEL_COUNT; ++i) {
    if (name && std::strcmp(name, kGeneralModels[i]) == 0) return i;
  }
  throw std::runtime_error(std::string("Unknown continuous-time model '") + (name ? name : "") +
                           "'");
}

void check_general_pars(int model, const double* row) {
  if (model == odem::MODEL_MALARIA && row[3] != odem::malaria::n_rates) {
    throw std::runtime_error("'n_rates' must be " + std::to_string(odem::malaria::n_rates) +
                             " (the exposed-vector stages are fixed at build time)");
  }
  if (model == odem::MODEL_CHANGE) {
    const double n = row[1];
    if (!(n >= 0 && n <= odem::change::max_changes && std::floor(n) == n)) {
      throw std::runtime_error("'n' must be an integer between 0 and " +
                               std::to_string(odem::change::max_changes));
    }
  }
  if (model == odem::MODEL_LOGDELAY && !(row[2] > 0)) {
    throw std::runtime_error("'tau' must be positive");
  }
}

void check_general(dust2_ode_general* h, const char* action) {
  check_handle(h, TAG_ODE_GENERAL, "system");
  if (h->errors_pending) {
    throw std::runtime_error(std::string("Can't currently ") + action +
                             " this system: errors are pending");
  }
}

// active: [G] groups to run (the others keep their state, time and solver
// flags), or null for every group.
void general_run(dust2_ode_general* h, const double* targets, size_t n, double* out,
                 const std::vector<uint8_t>* active = nullptr) {
  hipStream_t s = h->stream;
  const size_t n_total = h->G * h->P;
  h->d_targets.alloc(std::max<size_t>(n, 1) * sizeof(double));
  if (n) to_device(h->d_targets, targets, n * sizeof(double), s);
  to_device(h->d_reinit, h->reinit.data(), h->G, s);
  if (active) {
    h->d_active.alloc(h->G);
    to_device(h->d_active, active->data(), h->G, s);
  }
  if (out) h->d_out.alloc(std::max<size_t>(n, 1) * h->S * n_total * sizeof(double));
  hip_check(hipMemsetAsync(h->d_counters.get(), 0, h->d_counters.bytes(), s), "memset");
  gen_launch a{};
  a.model = h->model;
  a.pars = h->d_pars.get<double>();
  a.n_pars = h->info.n_pars;
  a.n_particles = static_cast<int>(h->P);
  a.n_total = static_cast<int>(n_total);
  a.reinit = h->d_reinit.get<uint8_t>();
  a.active = active ? h->d_active.get<uint8_t>() : nullptr;
  a.time = h->time;
  a.targets = h->d_targets.get<double>();
  a.n_targets = static_cast<int>(n);
  a.control = h->control;
  a.state = h->d_state.get<double>();
  a.internals = h->d_internals.get<double>();
  a.words = h->words;
  a.out = out ? h->d_out.get<double>() : nullptr;
  a.record = ode_record{h->record_cap ? h->d_step_times.get<double>() : nullptr,
                        h->d_step_count.get<uint32_t>(), h->record_cap};
  a.events = h->event_cap ? h->d_events.get<double>() : nullptr;
  a.event_cap = h->event_cap;
  a.history = h->history_cap ? h->d_history.get<double>() : nullptr;
  a.history_cap = h->history_cap;
  a.save_history = h->save_history ? 1 : 0;
  a.dt = h->dt;
  a.rng = h->info.mixed ? h->d_rng.get<uint64_t>() : nullptr;
  a.deterministic = h->deterministic ? 1 : 0;
  a.errors = h->d_errors.get<error_record>();
  a.error_count = h->d_counters.get<uint32_t>();
  hip_check(launch_general_system(a, s), "general ODE system kernel");
  uint32_t count = 0;
  to_host(&count, h->d_counters, sizeof(count), s);
  // initialise_solver_ clears the flags of the groups it ran (system.hpp:444-482)
  for (size_t g = 0; g < h->G; ++g) {
    if (!active || (*active)[g]) h->reinit[g] = 0;
  }
  if (count != 0) {
    std::vector<error_record> rec(n_total);
    to_host(rec.data(), h->d_errors, rec.size() * sizeof(error_record), s);
    hip_check(hipMemsetAsync(h->d_errors.get(), 0, h->d_errors.bytes(), s), "memset");
    hip_check(hipStreamSynchronize(s), "sync");
    h->errors_pending = true;
    throw std::runtime_error(error_report(rec, count));
  }
  if (n) h->time = targets[n - 1];
  if (out) {
    const size_t block = h->S * n_total;
    std::vector<double> tmp(n * block);
    to_host(tmp.data(), h->d_out, tmp.size() * sizeof(double), s);
    for (size_t q = 0; q < n; ++q) {
      soa_to_aos(tmp.data() + q * block, out + q * block, 1, n_total, h->S, true);
    }
  }
}

// Per-particle rows of a device array of n_total rows of `row` bytes,
// permuted within groups: row to <- row (group start + index[to]).
void permute_rows(dbuf& d, size_t G, size_t P, size_t row, const uint32_t* index, hipStream_t s) {
  if (row == 0 || d.get() == nullptr) return;
  const size_t n_total = G * P;
  std::vector<unsigned char> a(n_total * row), b(n_total * row);
  to_host(a.data(), d, a.size(), s);
  for (size_t g = 0; g < G; ++g) {
    for (size_t p = 0; p < P; ++p) {
      const size_t to = g * P + p, from = g * P + index[to];
      std::memcpy(b.data() + to * row, a.data() + from * row, row);
    }
  }
  to_device(d, b.data(), b.size(), s);
  hip_check(hipStreamSynchronize(s), "sync");
}

}  // namespace

extern "C" {

int dust2_system_ode_model(const char* model, size_t* sizes, char* err, size_t errlen) {
  try {
    const int id = general_model_id(model);
    const odem::model_info m = odem::info(id);
    sizes[0] = static_cast<size_t>(m.n_var);
    sizes[1] = static_cast<size_t>(m.n_special);
    sizes[2] = static_cast<size_t>(m.n_output);
    sizes[3] = static_cast<size_t>(m.n_pars);
    sizes[4] = static_cast<size_t>(m.max_events);
    sizes[5] = static_cast<size_t>(m.has_delays ? m.n_hist : m.n_var + m.n_special);
    sizes[6] = m.has_delays ? 1u : 0u;
    sizes[7] = m.mixed ? 1u : 0u;
    sizes[8] = static_cast<size_t>(m.n_data);
    return 0;
  } catch (const std::exception& e) {
    return fail(e, err, errlen);
  }
}

int dust2_system_ode_alloc(const char* model, const double* pars, size_t n_groups,
                           size_t n_particles, double time, double dt, const double* control,
                           const double* critical_times, size_t n_critical,
                           size_t record_step_times, size_t save_history, size_t history_cap,
                           size_t event_cap, const uint64_t* seed_state, size_t n_seed_streams,
                           uint64_t seed, int deterministic, dust2_ode_general** out, char* err,
                           size_t errlen) {
  try {
    check_sizes(n_groups, n_particles);
    auto h = std::make_unique<dust2_ode_general>();
    h->model = general_model_id(model);
    h->info = odem::info(h->model);
    if (dt != 0) {
      // dt is for mixed-time models only (R/interface.R:600-620)
      if (!h->info.mixed) throw std::runtime_error("Can't use 'dt' with continuous-time systems");
      check_dt(dt);
    }
    h->dt = dt;
    h->deterministic = deterministic != 0;
    const size_t np = static_cast<size_t>(h->info.n_pars);
    for (size_t g = 0; g < n_groups; ++g) check_general_pars(h->model, pars + g * np);
    h->G = n_groups;
    h->P = n_particles;
    h->S = static_cast<size_t>(h->info.n_var + h->info.n_special + h->info.n_output);
    h->words = gen_words(h->model);
    h->hist_words = gen_history_words(h->model);
    h->time = time;
    h->pars.assign(pars, pars + n_groups * np);
    h->reinit.assign(n_groups, 1);
    h->control = ode_control_from(control);
    h->record_cap = static_cast<uint32_t>(std::min<size_t>(record_step_times, 1u << 20));
    h->event_cap =
        h->info.max_events > 0 ? static_cast<uint32_t>(std::min<size_t>(event_cap, 1u << 20)) : 0;
    // a model with delays always keeps its history (system.hpp:414-440)
    h->save_history = save_history != 0 || h->info.has_delays;
    h->history_cap =
        h->save_history ? static_cast<uint32_t>(std::min<size_t>(std::max<size_t>(history_cap, 1),
                                                                 1u << 20))
                        : 0;
    const size_t n_total = n_groups * n_particles;
    if (static_cast<double>(n_total) * h->history_cap * h->hist_words * 8 > 64e9) {
      throw std::runtime_error("The history ring would exceed 64 GB: lower 'history_cap'");
    }
    hip_check(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking), "hipStreamCreate");
    h->d_pars.alloc(std::max<size_t>(h->pars.size(), 1) * sizeof(double));
    to_device(h->d_pars, pars, h->pars.size() * sizeof(double), h->stream);
    h->d_state.alloc(h->S * n_total * sizeof(double));
    h->d_internals.alloc(n_total * h->words * sizeof(double));
    h->d_reinit.alloc(n_groups);
    h->d_errors.alloc(n_total * sizeof(error_record));
    h->d_counters.alloc(4 * sizeof(uint32_t));
    h->d_step_count.alloc(n_total * sizeof(uint32_t));
    if (h->record_cap) h->d_step_times.alloc(n_total * h->record_cap * sizeof(double));
    if (h->event_cap) h->d_events.alloc(n_total * h->event_cap * 3 * sizeof(double));
    if (h->history_cap) {
      h->d_history.alloc(n_total * h->history_cap * h->hist_words * sizeof(double));
    }
    hip_check(hipMemsetAsync(h->d_state.get(), 0, h->d_state.bytes(), h->stream), "memset");
    hip_check(hipMemsetAsync(h->d_internals.get(), 0, h->d_internals.bytes(), h->stream),
              "memset");
    hip_check(hipMemsetAsync(h->d_errors.get(), 0, h->d_errors.bytes(), h->stream), "memset");
    hip_check(hipMemsetAsync(h->d_step_count.get(), 0, h->d_step_count.bytes(), h->stream),
              "memset");
    if (h->info.mixed) {
      // one stream per particle, as the system's prng (system.hpp:361)
      h->d_rng.alloc(n_total * 4 * sizeof(uint64_t));
      expand_streams(seed_state, n_seed_streams, seed, n_total, h->d_rng.get<uint64_t>(),
                     h->stream);
    }
    set_critical_times(h->control, h->d_critical, critical_times, n_critical, h->stream);
    *out = h.release();
    return 0;
  } catch (const std::exception& e) {
    return fail(e, err, errlen);
  }
}

int dust2_system_ode_rng_state(dust2_ode_general* h, uint64_t* out, char* err, size_t errlen) {
  try {
    check_handle(h, TAG_ODE_GENERAL, "system");
    if (!h->info.mixed) throw std::runtime_error("This system has no random number streams");
    to_host(out, h->d_rng, h->G * h->P * 4 * sizeof(uint64_t), h->stream);
    return 0;
  } catch (const std::exception& e) {
    return fail(e, err, errlen);
  }
}

int dust2_system_ode_set_rng_state(dust2_ode_general* h, const uint64_t* state, char* err,
                                   size_t errlen) {
  try {
    check_handle(h, TAG_ODE_GENERAL, "system");
    if (!h->info.mixed) throw std::runtime_error("This system has no random number streams");
    to_device(h->d_rng, state, h->G * h->P * 4 * sizeof(uint64_t), h->stream);
    hip_check(hipStreamSynchronize(h->stream), "sync");
    return 0;
  } catch (const std::exception& e) {
    return fail(e, err, errlen);
  }
}

int dust2_system_ode_compare_data(dust2_ode_general* h, const double* data,
                                  size_t n_groups_data, double* out, char* err, size_t errlen) {
  try {
    check_general(h, "compare data for");
    if (h->info.n_data == 0) throw std::runtime_error("This model has no compare_data");
    if (n_groups_data != 1 && n_groups_data != h->G) {
      throw std::runtime_error("Expected data for one group or every group");
    }
    const size_t n_total = h->G * h->P;
    const size_t nd = n_groups_data * static_cast<size_t>(h->info.n_data);
    h->d_data.alloc(nd * sizeof(double));
    to_device(h->d_data, data, nd * sizeof(double), h->stream);
    h->d_ll.alloc(n_total * sizeof(double));
    gen_compare_launch a{};
    a.model = h->model;
    a.pars = h->d_pars.get<double>();
    a.n_pars = h->info.n_pars;
    a.n_particles = static_cast<int>(h->P);
    a.n_total = static_cast<int>(n_total);
    a.time = h->time;
    a.state = h->d_state.get<double>();
    a.data = h->d_data.get<double>();
    a.shared_data = n_groups_data == 1 ? 1 : 0;
    a.out = h->d_ll.get<double>();
    hip_check(launch_general_compare(a, h->stream), "compare kernel");
    to_host(out, h->d_ll, n_total * sizeof(double), h->stream);
    return 0;
  } catch (const std::exception& e) {
    return fail(e, err, errlen);
  }
}

int dust2_system_ode_run_to_time(dust2_ode_general* h, double time, char* err, size_t errlen) {
  try {
    check_general(h, "run");
    if (time < h->time) {
      char buf[256];
      std::snprintf(buf, sizeof buf, "Can't run to time %f, system already at time %f", time,
                    h->time);
      throw std::runtime_error(buf);
    }
    general_run(h, &time, 1, nullptr);
    return 0;
  } catch (const std::exception& e) {
    return fail(e, err, errlen);
  }
}

int dust2_system_ode_simulate(dust2_ode_general* h, const double* times, size_t n_times,
                              double* out, char* err, size_t errlen) {
  try {
    check_general(h, "simulate");
    double prev = h->time;
    for (size_t i = 0; i < n_times; ++i) {
      if (times[i] < prev) {
        throw std::runtime_error("Expected 'times' to be increasing, starting at the "
                                 "system's time");
      }
      prev = times[i];
    }
    general_run(h, times, n_times, out);
    return 0;
  } catch (const std::exception& e) {
    return fail(e, err, errlen);
  }
}

int dust2_system_ode_state(dust2_ode_general* h, double* out, char* err, size_t errlen) {
  try {
    check_general(h, "get state from");
    std::vector<double> tmp(h->S * h->G * h->P);
    to_host(tmp.data(), h->d_state, tmp.size() * sizeof(double), h->stream);
    soa_to_aos(tmp.data(), out, 1, h->G * h->P, h->S, true);
    return 0;
  } catch (const std::exception& e) {
    return fail(e, err, errlen);
  }
}

int dust2_system_ode_set_state(dust2_ode_general* h, const double* state, char* err,
                               size_t errlen) {
  try {
    check_handle(h, TAG_ODE_GENERAL, "system");  // allowed while errors are pending
    std::vector<double> tmp(h->S * h->G * h->P);
    aos_to_soa(state, tmp.data(), 1, h->G * h->P, h->S, true);
    to_device(h->d_state, tmp.data(), tmp.size() * sizeof(double), h->stream);
    hip_check(hipStreamSynchronize(h->stream), "sync");
    std::fill(h->reinit.begin(), h->reinit.end(), 1);
    h->errors_pending = false;
    return 0;
  } catch (const std::exception& e) {
    return fail(e, err, errlen);
  }
}

int dust2_system_ode_set_state_initial(dust2_ode_general* h, char* err, size_t errlen) {
  try {
    check_handle(h, TAG_ODE_GENERAL, "system");  // allowed while errors are pending
    std::vector<double> st(h->G * h->P * h->S, 0.0);
    const size_t np = static_cast<size_t>(h->info.n_pars);
    for (size_t g = 0; g < h->G; ++g) {
      const double* p = h->pars.data() + g * np;
      for (size_t q = 0; q < h->P; ++q) {
        double* y = st.data() + (g * h->P + q) * h->S;
        switch (h->model) {
          case odem::MODEL_BOUNCE: odem::bounce::initial(h->time, p, y); break;
          case odem::MODEL_CHANGE: odem::change::initial(h->time, p, y); break;
          case odem::MODEL_EVENTZERO: odem::eventzero::initial(h->time, p, y); break;
          case odem::MODEL_SPECIALEVENT: odem::specialevent::initial(h->time, p, y); break;
          case odem::MODEL_SIRDELAY: odem::sirdelay::initial(h->time, p, y); break;
          case odem::MODEL_LOGDELAY: odem::logdelay::initial(h->time, p, y); break;
          case odem::MODEL_SPECIAL: odem::special::initial(h->time, p, y); break;
          case odem::MODEL_MALARIA: odem::malaria::initial(h->time, p, y); break;
          default: break;
        }
      }
    }
    h->errors_pending = false;
    return dust2_system_ode_set_state(h, st.data(), err, errlen);
  } catch (const std::exception& e) {
    return fail(e, err, errlen);
  }
}

int dust2_system_ode_time(dust2_ode_general* h, double* out, char* err, size_t errlen) {
  try {
    check_general(h, "get time from");
    *out = h->time;
    return 0;
  } catch (const std::exception& e) {
    return fail(e, err, errlen);
  }
}

int dust2_system_ode_set_time(dust2_ode_general* h, double time, char* err, size_t errlen) {
  try {
    check_general(h, "set time for");
    h->time = time;
    std::fill(h->reinit.begin(), h->reinit.end(), 1);
    return 0;
  } catch (const std::exception& e) {
    return fail(e, err, errlen);
  }
}

int dust2_system_ode_update_pars(dust2_ode_general* h, const double* pars, char* err,
                                 size_t errlen) {
  try {
    check_handle(h, TAG_ODE_GENERAL, "system");  // allowed while errors are pending
    const size_t np = static_cast<size_t>(h->info.n_pars);
    for (size_t g = 0; g < h->G; ++g) check_general_pars(h->model, pars + g * np);
    h->pars.assign(pars, pars + h->G * np);
    to_device(h->d_pars, pars, h->pars.size() * sizeof(double), h->stream);
    hip_check(hipStreamSynchronize(h->stream), "sync");
    // update_shared marks the group's solver for initialisation
    // (dust2/continuous/system.hpp:256-265)
    std::fill(h->reinit.begin(), h->reinit.end(), 1);
    return 0;
  } catch (const std::exception& e) {
    return fail(e, err, errlen);
  }
}

int dust2_system_ode_reorder(dust2_ode_general* h, const uint32_t* index, char* err,
                             size_t errlen) {
  try {
    check_general(h, "reorder");
    const size_t n_total = h->G * h->P;
    for (size_t k = 0; k < n_total; ++k) {
      if (index[k] >= h->P) throw std::runtime_error("reorder index out of range");
    }
    // state and every piece of solver internals move together
    // (dust2/continuous/system.hpp:172-199)
    std::vector<double> st(h->S * n_total), st2(h->S * n_total);
    to_host(st.data(), h->d_state, st.size() * sizeof(double), h->stream);
    for (size_t g = 0; g < h->G; ++g) {
      for (size_t p = 0; p < h->P; ++p) {
        const size_t to = g * h->P + p, from = g * h->P + index[to];
        for (size_t c = 0; c < h->S; ++c) st2[c * n_total + to] = st[c * n_total + from];
      }
    }
    to_device(h->d_state, st2.data(), st2.size() * sizeof(double), h->stream);
    permute_rows(h->d_internals, h->G, h->P, h->words * sizeof(double), index, h->stream);
    permute_rows(h->d_step_count, h->G, h->P, sizeof(uint32_t), index, h->stream);
    permute_rows(h->d_step_times, h->G, h->P, h->record_cap * sizeof(double), index, h->stream);
    permute_rows(h->d_events, h->G, h->P, h->event_cap * 3 * sizeof(double), index, h->stream);
    permute_rows(h->d_history, h->G, h->P,
                 static_cast<size_t>(h->history_cap) * h->hist_words * sizeof(double), index,
                 h->stream);
    return 0;
  } catch (const std::exception& e) {
    return fail(e, err, errlen);
  }
}

int dust2_system_ode_internals(dust2_ode_general* h, double* stats, double* step_times,
                               size_t* n_step_times, char* err, size_t errlen) {
  try {
    check_general(h, "get internals from");
    const size_t n_total = h->G * h->P;
    const odem::model_info& m = h->info;
    const int ny = m.n_var + m.n_special;
    const gen_layout L = make_gen_layout(m.n_var, ny, m.max_events > 0 ? m.max_events : 1,
                                         m.has_delays ? m.n_hist : ny);
    std::vector<double> w(n_total * h->words);
    to_host(w.data(), h->d_internals, w.size() * sizeof(double), h->stream);
    for (size_t k = 0; k < n_total; ++k) {
      const double* r = w.data() + k * h->words;
      stats[4 * k + 0] = r[L.n_steps];
      stats[4 * k + 1] = r[L.n_accepted];
      stats[4 * k + 2] = r[L.n_rejected];
      stats[4 * k + 3] = r[L.step_size];
    }
    if (n_step_times) {
      std::vector<uint32_t> cnt(n_total);
      to_host(cnt.data(), h->d_step_count, cnt.size() * sizeof(uint32_t), h->stream);
      for (size_t k = 0; k < n_total; ++k) n_step_times[k] = h->record_cap ? cnt[k] : 0;
      if (step_times && h->record_cap) {
        to_host(step_times, h->d_step_times, n_total * h->record_cap * sizeof(double),
                h->stream);
      }
    }
    return 0;
  } catch (const std::exception& e) {
    return fail(e, err, errlen);
  }
}

int dust2_system_ode_events(dust2_ode_general* h, size_t* counts, double* records, char* err,
                            size_t errlen) {
  try {
    check_general(h, "get internals from");
    const size_t n_total = h->G * h->P;
    const odem::model_info& m = h->info;
    const int ny = m.n_var + m.n_special;
    const gen_layout L = make_gen_layout(m.n_var, ny, m.max_events > 0 ? m.max_events : 1,
                                         m.has_delays ? m.n_hist : ny);
    std::vector<double> w(n_total * h->words);
    to_host(w.data(), h->d_internals, w.size() * sizeof(double), h->stream);
    for (size_t k = 0; k < n_total; ++k) {
      counts[k] = static_cast<size_t>(w[k * h->words + L.event_count]);
    }
    if (records && h->event_cap) {
      to_host(records, h->d_events, n_total * h->event_cap * 3 * sizeof(double), h->stream);
    }
    return 0;
  } catch (const std::exception& e) {
    return fail(e, err, errlen);
  }
}

int dust2_system_ode_history(dust2_ode_general* h, size_t* counts, double* entries, char* err,
                             size_t errlen) {
  try {
    check_general(h, "get internals from");
    const size_t n_total = h->G * h->P;
    const odem::model_info& m = h->info;
    const int ny = m.n_var + m.n_special;
    const gen_layout L = make_gen_layout(m.n_var, ny, m.max_events > 0 ? m.max_events : 1,
                                         m.has_delays ? m.n_hist : ny);
    std::vector<double> w(n_total * h->words);
    to_host(w.data(), h->d_internals, w.size() * sizeof(double), h->stream);
    const size_t cap = h->history_cap, hw = static_cast<size_t>(h->hist_words);
    std::vector<double> ring(cap ? n_total * cap * hw : 0);
    if (cap) to_host(ring.data(), h->d_history, ring.size() * sizeof(double), h->stream);
    for (size_t k = 0; k < n_total; ++k) {
      const size_t count = static_cast<size_t>(w[k * h->words + L.hist_count]);
      counts[k] = count;
      if (!entries || cap == 0) continue;
      // the kept steps, oldest first
      const size_t kept = std::min(count, cap), first = count - kept;
      for (size_t j = 0; j < kept; ++j) {
        const double* src = ring.data() + (k * cap + (first + j) % cap) * hw;
        std::copy(src, src + hw, entries + (k * cap + j) * hw);
      }
    }
    return 0;
  } catch (const std::exception& e) {
    return fail(e, err, errlen);
  }
}

void dust2_system_ode_free(dust2_ode_general* h) {
  if (handle_ok(h, TAG_ODE_GENERAL)) delete h;
}

}  // extern "C"

// ============================================================================
// Particle filter over a general continuous / mixed-time model
// (dust2/filter.hpp:42-175 over dust_continuous)
// ============================================================================
struct dust2_filter_ode {
  uint32_t tag = TAG_FILTER_ODE;
  dust2_ode_general* sys = nullptr;   // owned
  size_t G = 0, P = 0, n_time = 0, n_groups_data = 0;
  double time_start = 0;
  std::vector<double> times;
  dbuf d_data, d_frng, d_ll, d_ll_step, d_idx, d_state2, d_words2, d_traj, d_ancestry,
      d_fallback, d_groups, d_out, d_run_groups, d_states, d_slot, d_select, d_todo;
  std::vector<uint32_t> last_groups;  // groups of the last run
  std::vector<uint8_t> select_valid;  // chosen particle drawn this run (per group)
  saved_run saved;                    // trajectories and snapshots of the last run
  ~dust2_filter_ode() {
    if (sys) dust2_system_ode_free(sys);
  }
};

namespace {

void check_filter_ode(dust2_filter_ode* h) { check_handle(h, TAG_FILTER_ODE, "filter"); }

// index_group (0-based, or null for every group) as a run list and a mask
void run_groups(const size_t* index_group, size_t n_index, size_t G,
                std::vector<uint32_t>& groups, std::vector<uint8_t>& active) {
  groups.clear();
  active.assign(G, 0);
  if (index_group == nullptr) {
    for (size_t g = 0; g < G; ++g) groups.push_back(static_cast<uint32_t>(g));
    std::fill(active.begin(), active.end(), 1);
    return;
  }
  for (size_t i = 0; i < n_index; ++i) {
    if (index_group[i] >= G) throw std::runtime_error("'index_group' out of range");
    if (active[index_group[i]]) throw std::runtime_error("'index_group' has duplicates");
    active[index_group[i]] = 1;
    groups.push_back(static_cast<uint32_t>(index_group[i]));
  }
}

}  // namespace

extern "C" {

int dust2_filter_ode_alloc(const char* model, const double* pars, size_t n_groups,
                           size_t n_particles, double time_start, const double* time,
                           size_t n_time, const double* data, size_t n_groups_data, double dt,
                           const double* control, const uint64_t* seed_state,
                           size_t n_seed_streams, uint64_t seed, dust2_filter_ode** out,
                           char* err, size_t errlen) {
  try {
    check_sizes(n_groups, n_particles);
    const int id = general_model_id(model);
    const odem::model_info info = odem::info(id);
    if (info.n_data == 0) throw std::runtime_error("This model has no compare_data");
    if (info.has_delays) throw std::runtime_error("Filters over models with delays are not built");
    if (!info.mixed) {
      throw std::runtime_error("Can't use 'dust_filter_create()' with continuous-time models");
    }
    if (n_particles > 8192) throw std::runtime_error("At most 8192 particles per group");
    if ((kScratchDoubles + n_particles) * sizeof(double) >
        static_cast<size_t>(max_lds_per_block())) {
      throw std::runtime_error("Too many particles per group for the resampling workgroup");
    }
    check_time_sequence(time_start, time, n_time);
    if (n_groups_data != 1 && n_groups_data != n_groups) {
      throw std::runtime_error("Expected data for one group or every group");
    }
    auto h = std::make_unique<dust2_filter_ode>();
    h->G = n_groups;
    h->P = n_particles;
    h->n_time = n_time;
    h->n_groups_data = n_groups_data;
    h->time_start = time_start;
    h->times.assign(time, time + n_time);
    const size_t G = n_groups, P = n_particles;
    // streams [2 filter | P particle] per group (dust2/r/discrete/filter.hpp:63-74)
    std::vector<uint64_t> all(G * (P + 2) * 4), part(G * P * 4), filt(G * 8);
    {
      hipStream_t s = nullptr;
      hip_check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate");
      dbuf tmp(all.size() * sizeof(uint64_t));
      expand_streams(seed_state, n_seed_streams, seed, G * (P + 2), tmp.get<uint64_t>(), s);
      to_host(all.data(), tmp, all.size() * sizeof(uint64_t), s);
      (void)hipStreamDestroy(s);
    }
    for (size_t g = 0; g < G; ++g) {
      const uint64_t* blk = all.data() + g * (P + 2) * 4;
      std::copy(blk, blk + 8, filt.data() + g * 8);
      std::copy(blk + 8, blk + (P + 2) * 4, part.data() + g * P * 4);
    }
    char buf[1024];
    if (dust2_system_ode_alloc(model, pars, n_groups, n_particles, time_start, dt, control,
                               nullptr, 0, 0, 0, 0, 0, part.data(), G * P, 0, 0, &h->sys, buf,
                               sizeof buf) != 0) {
      throw std::runtime_error(buf);
    }
    hipStream_t s = h->sys->stream;
    const size_t nd = n_time * n_groups_data * static_cast<size_t>(info.n_data);
    h->d_data.alloc(nd * sizeof(double));
    to_device(h->d_data, data, nd * sizeof(double), s);
    h->d_frng.alloc(filt.size() * sizeof(uint64_t));
    to_device(h->d_frng, filt.data(), filt.size() * sizeof(uint64_t), s);
    h->d_ll.alloc(G * sizeof(double));
    h->d_ll_step.alloc(G * P * sizeof(double));
    h->d_idx.alloc(G * P * sizeof(uint32_t));
    h->d_state2.alloc(h->sys->d_state.bytes());
    h->d_words2.alloc(h->sys->d_internals.bytes());
    h->d_fallback.alloc(sizeof(uint32_t));
    h->d_select.alloc(G * sizeof(uint32_t));
    h->select_valid.assign(G, 0);
    hip_check(hipStreamSynchronize(s), "sync");
    *out = h.release();
    return 0;
  } catch (const std::exception& e) {
    return fail(e, err, errlen);
  }
}

// filter<T>::run (dust2/filter.hpp:42-104) with reset (:156-175): every
// group reset to time_start (initial: [G][P][S] or null for the model's
// own), then for the groups of index_group (null: all) per data time
// run_to_time, compare_data, log-sum-exp, one uniform per group and
// systematic resampling, then reorder.  ll_out[n_index] in index order.
int dust2_filter_ode_run(dust2_filter_ode* h, const double* initial, int save_trajectories_in,
                         const double* save_snapshots, size_t n_snapshots,
                         const size_t* index_state, size_t n_index_state,
                         const size_t* index_group, size_t n_index, double* ll_out, char* err,
                         size_t errlen) {
  try {
    check_filter_ode(h);
    dust2_ode_general* sys = h->sys;
    hipStream_t s = sys->stream;
    const size_t G = h->G, P = h->P, n_total = G * P;
    char buf[1024];
    std::vector<uint32_t> groups;
    std::vector<uint8_t> active;
    run_groups(index_group, n_index, G, groups, active);
    saved_run req = check_save(save_trajectories_in, save_snapshots, n_snapshots, index_state,
                               n_index_state, sys->S, h->times);
    // snapshots follow the ancestry of the saved states (dust2/filter.hpp:47-49)
    const bool save_trajectories = req.any();
    const bool all = groups.size() == G;
    h->d_run_groups.alloc(G * sizeof(uint32_t));
    to_device(h->d_run_groups, groups.data(), groups.size() * sizeof(uint32_t), s);
    // reset: set_time(time_start), then the initial state
    sys->errors_pending = false;
    sys->time = h->time_start;
    std::fill(sys->reinit.begin(), sys->reinit.end(), 1);
    const int rc = initial ? dust2_system_ode_set_state(sys, initial, buf, sizeof buf)
                           : dust2_system_ode_set_state_initial(sys, buf, sizeof buf);
    if (rc != 0) throw std::runtime_error(buf);
    hip_check(hipMemsetAsync(h->d_ll.get(), 0, h->d_ll.bytes(), s), "memset");
    hip_check(hipMemsetAsync(h->d_fallback.get(), 0, h->d_fallback.bytes(), s), "memset");
    h->saved = saved_run{};
    // a new run forgets the chosen particles (dust2/filter.hpp:173)
    std::fill(h->select_valid.begin(), h->select_valid.end(), 0);
    h->last_groups.clear();
    if (save_trajectories) {
      h->d_traj.alloc(G * h->n_time * sys->S * P * sizeof(double));
      h->d_ancestry.alloc(G * h->n_time * P * sizeof(uint32_t));
    }
    const int nd = sys->info.n_data;
    for (size_t t = 0; t < h->n_time; ++t) {
      general_run(sys, &h->times[t], 1, nullptr, all ? nullptr : &active);
      gen_compare_launch c{};
      c.model = sys->model;
      c.pars = sys->d_pars.get<double>();
      c.n_pars = sys->info.n_pars;
      c.n_particles = static_cast<int>(P);
      c.n_total = static_cast<int>(n_total);
      c.time = sys->time;
      c.state = sys->d_state.get<double>();
      c.data = h->d_data.get<double>() + t * h->n_groups_data * nd;
      c.shared_data = h->n_groups_data == 1 ? 1 : 0;
      c.out = h->d_ll_step.get<double>();
      c.active = all ? nullptr : sys->d_active.get<uint8_t>();
      hip_check(launch_general_compare(c, s), "compare kernel");
      gen_resample_launch r{};
      r.ll_step = h->d_ll_step.get<double>();
      r.filter_rng = h->d_frng.get<uint64_t>();
      r.ll = h->d_ll.get<double>();
      r.idx = h->d_idx.get<uint32_t>();
      r.groups = h->d_run_groups.get<uint32_t>();
      r.n_groups = static_cast<int>(groups.size());
      r.n_particles = static_cast<int>(P);
      r.fallback_count = h->d_fallback.get<uint32_t>();
      hip_check(launch_general_resample(r, s), "resample kernel");
      gen_reorder_launch o{};
      o.idx = h->d_idx.get<uint32_t>();
      o.state_in = sys->d_state.get<double>();
      o.state_out = h->d_state2.get<double>();
      o.S = static_cast<int>(sys->S);
      o.words_in = sys->d_internals.get<double>();
      o.words_out = h->d_words2.get<double>();
      o.words = sys->words;
      o.n_particles = static_cast<int>(P);
      o.n_total = static_cast<int>(n_total);
      o.traj = save_trajectories ? h->d_traj.get<double>() : nullptr;
      o.ancestry = save_trajectories ? h->d_ancestry.get<uint32_t>() : nullptr;
      o.n_time = static_cast<int>(h->n_time);
      o.t = static_cast<int>(t);
      o.active = all ? nullptr : sys->d_active.get<uint8_t>();
      hip_check(launch_general_reorder(o, s), "reorder kernel");
      sys->d_state.swap(h->d_state2);
      sys->d_internals.swap(h->d_words2);
    }
    std::vector<double> ll(G);
    to_host(ll.data(), h->d_ll, G * sizeof(double), s);
    for (size_t i = 0; i < groups.size(); ++i) ll_out[i] = ll[groups[i]];
    h->last_groups = groups;
    h->saved = std::move(req);
    return 0;
  } catch (const std::exception& e) {
    return fail(e, err, errlen);
  }
}

int dust2_filter_ode_update_pars(dust2_filter_ode* h, const double* pars, char* err,
                                 size_t errlen) {
  try {
    check_filter_ode(h);
    return dust2_system_ode_update_pars(h->sys, pars, err, errlen);
  } catch (const std::exception& e) {
    return fail(e, err, errlen);
  }
}

int dust2_filter_ode_last_state(dust2_filter_ode* h, double* out, char* err, size_t errlen) {
  try {
    check_filter_ode(h);
    h->sys->errors_pending = false;
    return dust2_system_ode_state(h->sys, out, err, errlen);
  } catch (const std::exception& e) {
    return fail(e, err, errlen);
  }
}

}  // extern "C"

namespace {
// filter<T>::select_random_particle (dust2/filter.hpp:130-138) for the groups
// of the last run: a group's particle is drawn once per run, from its second
// filter stream, the first time any export asks for it.
void filter_ode_select(dust2_filter_ode* h) {
  hipStream_t s = h->sys->stream;
  std::vector<uint32_t> todo;
  for (uint32_t g : h->last_groups) {
    if (!h->select_valid[g]) todo.push_back(g);
  }
  if (todo.empty()) return;
  h->d_todo.alloc(todo.size() * sizeof(uint32_t));
  to_device(h->d_todo, todo.data(), todo.size() * sizeof(uint32_t), s);
  hip_check(launch_select_particle(h->d_frng.get<uint64_t>(), h->d_todo.get<uint32_t>(),
                                   static_cast<int>(todo.size()), static_cast<int>(h->P), 8, 4,
                                   1, h->d_select.get<uint32_t>(), s),
            "select particle kernel");
  hip_check(hipStreamSynchronize(s), "sync");
  for (uint32_t g : todo) h->select_valid[g] = 1;
}

// Trajectories (snapshots = false: the states of index_state at every data
// time) or snapshots (the whole state at the snapshot times) following the
// ancestry (dust2/trajectories.hpp:128-283) of the groups of the last run:
// out [times][n_groups_run][P, or 1 when select][states].
void filter_ode_export(dust2_filter_ode* h, bool snapshots, bool select, double* out) {
  hipStream_t s = h->sys->stream;
  const size_t P = h->P, S = h->sys->S, T = h->n_time;
  const std::vector<uint32_t>& groups = h->last_groups;
  const size_t G = groups.size();
  const std::vector<uint32_t>& states = h->saved.states;
  const bool subset = !snapshots && !states.empty();
  const size_t s_out = subset ? states.size() : S;
  const size_t rows = snapshots ? h->saved.n_snap : T;
  h->d_groups.alloc(G * sizeof(uint32_t));
  to_device(h->d_groups, groups.data(), G * sizeof(uint32_t), s);
  if (subset) {
    h->d_states.alloc(s_out * sizeof(uint32_t));
    to_device(h->d_states, states.data(), s_out * sizeof(uint32_t), s);
  }
  if (snapshots) {
    h->d_slot.alloc(T * sizeof(int32_t));
    to_device(h->d_slot, h->saved.slot.data(), T * sizeof(int32_t), s);
  }
  if (select) filter_ode_select(h);
  const size_t bytes = rows * G * (select ? 1 : P) * s_out * sizeof(double);
  h->d_out.alloc(bytes);
  hip_check(launch_lineage(h->d_traj.get(), h->d_ancestry.get<uint32_t>(),
                           h->d_groups.get<uint32_t>(), static_cast<int>(G),
                           static_cast<int>(T), static_cast<int>(S), static_cast<int>(P),
                           select ? h->d_select.get<uint32_t>() : nullptr,
                           snapshots ? h->d_slot.get<int32_t>() : nullptr,
                           subset ? h->d_states.get<uint32_t>() : nullptr,
                           static_cast<int>(s_out), true, h->d_out.get<double>(), s),
            "lineage kernel");
  to_host(out, h->d_out, bytes, s);
}
}  // namespace

extern "C" {

int dust2_filter_ode_last_trajectories(dust2_filter_ode* h, int select_random_particle,
                                       double* out, char* err, size_t errlen) {
  try {
    check_filter_ode(h);
    if (!h->saved.traj) throw std::runtime_error("Trajectories are not current");
    filter_ode_export(h, false, select_random_particle != 0, out);
    return 0;
  } catch (const std::exception& e) {
    return fail(e, err, errlen);
  }
}

int dust2_filter_ode_last_snapshots(dust2_filter_ode* h, int select_random_particle,
                                    double* out, char* err, size_t errlen) {
  try {
    check_filter_ode(h);
    if (h->saved.n_snap == 0) throw std::runtime_error("Snapshots are not current");
    filter_ode_export(h, true, select_random_particle != 0, out);
    return 0;
  } catch (const std::exception& e) {
    return fail(e, err, errlen);
  }
}

int dust2_filter_ode_last_state_selected(dust2_filter_ode* h, double* out, char* err,
                                         size_t errlen) {
  try {
    check_filter_ode(h);
    if (h->last_groups.empty()) throw std::runtime_error("State is not current");
    filter_ode_select(h);
    const size_t P = h->P, S = h->sys->S;
    std::vector<uint32_t> sel(h->G);
    to_host(sel.data(), h->d_select, h->G * sizeof(uint32_t), h->sys->stream);
    std::vector<double> st(h->G * P * S);
    h->sys->errors_pending = false;
    char buf[1024];
    if (dust2_system_ode_state(h->sys, st.data(), buf, sizeof buf) != 0) {
      throw std::runtime_error(buf);
    }
    for (size_t i = 0; i < h->last_groups.size(); ++i) {
      const size_t g = h->last_groups[i];
      std::copy_n(st.data() + (g * P + sel[g]) * S, S, out + i * S);
    }
    return 0;
  } catch (const std::exception& e) {
    return fail(e, err, errlen);
  }
}

// the streams in the reference's export order: per group the 2 filter
// streams then the P particle streams (dust2/r/filter.hpp:182-207)
int dust2_filter_ode_rng_state(dust2_filter_ode* h, uint64_t* out, char* err, size_t errlen) {
  try {
    check_filter_ode(h);
    const size_t G = h->G, P = h->P;
    std::vector<uint64_t> filt(G * 8), part(G * P * 4);
    to_host(filt.data(), h->d_frng, filt.size() * sizeof(uint64_t), h->sys->stream);
    to_host(part.data(), h->sys->d_rng, part.size() * sizeof(uint64_t), h->sys->stream);
    for (size_t g = 0; g < G; ++g) {
      uint64_t* blk = out + g * (P + 2) * 4;
      std::copy(filt.data() + g * 8, filt.data() + g * 8 + 8, blk);
      std::copy(part.data() + g * P * 4, part.data() + (g + 1) * P * 4, blk + 8);
    }
    return 0;
  } catch (const std::exception& e) {
    return fail(e, err, errlen);
  }
}

int dust2_filter_ode_set_rng_state(dust2_filter_ode* h, const uint64_t* state, char* err,
                                   size_t errlen) {
  try {
    check_filter_ode(h);
    const size_t G = h->G, P = h->P;
    std::vector<uint64_t> filt(G * 8), part(G * P * 4);
    for (size_t g = 0; g < G; ++g) {
      const uint64_t* blk = state + g * (P + 2) * 4;
      std::copy(blk, blk + 8, filt.data() + g * 8);
      std::copy(blk + 8, blk + (P + 2) * 4, part.data() + g * P * 4);
    }
    to_device(h->d_frng, filt.data(), filt.size() * sizeof(uint64_t), h->sys->stream);
    to_device(h->sys->d_rng, part.data(), part.size() * sizeof(uint64_t), h->sys->stream);
    hip_check(hipStreamSynchronize(h->sys->stream), "sync");
    return 0;
  } catch (const std::exception& e) {
    return fail(e, err, errlen);
  }
}

void dust2_filter_ode_free(dust2_filter_ode* h) {
  if (handle_ok(h, TAG_FILTER_ODE)) delete h;
}

}  // extern "C"

// ============================================================================
// Deterministic likelihood ("unfilter") over a general model with data
// (dust2/unfilter.hpp:42-72, 162-186 over dust_continuous)
// ============================================================================
struct dust2_unfilter_ode {
  uint32_t tag = TAG_UNFILTER_ODE;
  dust2_ode_general* sys = nullptr;   // owned, deterministic
  size_t G = 0, P = 0, n_time = 0, n_groups_data = 0;
  double time_start = 0;
  std::vector<double> times;
  dbuf d_data, d_ll, d_traj;
  std::vector<uint32_t> last_groups;
  saved_run saved;                    // trajectories and snapshots of the last run
  bool has_run = false;
  ~dust2_unfilter_ode() {
    if (sys) dust2_system_ode_free(sys);
  }
};

extern "C" {

int dust2_unfilter_ode_alloc(const char* model, const double* pars, size_t n_groups,
                             size_t n_particles, double time_start, const double* time,
                             size_t n_time, const double* data, size_t n_groups_data, double dt,
                             const double* control, dust2_unfilter_ode** out, char* err,
                             size_t errlen) {
  try {
    check_sizes(n_groups, n_particles);
    const int id = general_model_id(model);
    const odem::model_info info = odem::info(id);
    if (info.n_data == 0) throw std::runtime_error("This model has no compare_data");
    check_time_sequence(time_start, time, n_time);
    if (n_groups_data != 1 && n_groups_data != n_groups) {
      throw std::runtime_error("Expected data for one group or every group");
    }
    auto h = std::make_unique<dust2_unfilter_ode>();
    h->G = n_groups;
    h->P = n_particles;
    h->n_time = n_time;
    h->n_groups_data = n_groups_data;
    h->time_start = time_start;
    h->times.assign(time, time + n_time);
    char buf[1024];
    // deterministic streams seeded with 42 (dust2/r/continuous/unfilter.hpp:38-42)
    if (dust2_system_ode_alloc(model, pars, n_groups, n_particles, time_start, dt, control,
                               nullptr, 0, 0, 0, 0, 0, nullptr, 0, 42, 1, &h->sys, buf,
                               sizeof buf) != 0) {
      throw std::runtime_error(buf);
    }
    hipStream_t s = h->sys->stream;
    const size_t nd = n_time * n_groups_data * static_cast<size_t>(info.n_data);
    h->d_data.alloc(nd * sizeof(double));
    to_device(h->d_data, data, nd * sizeof(double), s);
    h->d_ll.alloc(n_groups * n_particles * sizeof(double));
    hip_check(hipStreamSynchronize(s), "sync");
    *out = h.release();
    return 0;
  } catch (const std::exception& e) {
    return fail(e, err, errlen);
  }
}

// unfilter::run: every group reset to time_start (initial [G][P][S] or
// null), then for the groups of index_group (null: all) per data time
// run_to_time and compare_data, each particle's log-likelihood summed;
// ll_out[n_index][P].
int dust2_unfilter_ode_run(dust2_unfilter_ode* h, const double* initial, int save_trajectories_in,
                           const double* save_snapshots, size_t n_snapshots,
                           const size_t* index_state, size_t n_index_state,
                           const size_t* index_group, size_t n_index, double* ll_out,
                           char* err, size_t errlen) {
  try {
    check_handle(h, TAG_UNFILTER_ODE, "unfilter");
    dust2_ode_general* sys = h->sys;
    hipStream_t s = sys->stream;
    const size_t n_total = h->G * h->P;
    char buf[1024];
    std::vector<uint32_t> groups;
    std::vector<uint8_t> active;
    run_groups(index_group, n_index, h->G, groups, active);
    saved_run req = check_save(save_trajectories_in, save_snapshots, n_snapshots, index_state,
                               n_index_state, sys->S, h->times);
    const bool save_trajectories = req.any();
    const bool all = groups.size() == h->G;
    sys->errors_pending = false;
    sys->time = h->time_start;
    std::fill(sys->reinit.begin(), sys->reinit.end(), 1);
    const int rc = initial ? dust2_system_ode_set_state(sys, initial, buf, sizeof buf)
                           : dust2_system_ode_set_state_initial(sys, buf, sizeof buf);
    if (rc != 0) throw std::runtime_error(buf);
    hip_check(hipMemsetAsync(h->d_ll.get(), 0, h->d_ll.bytes(), s), "memset");
    h->saved = saved_run{};
    const size_t block = sys->S * n_total;
    if (save_trajectories) h->d_traj.alloc(h->n_time * block * sizeof(double));
    const int nd = sys->info.n_data;
    for (size_t t = 0; t < h->n_time; ++t) {
      general_run(sys, &h->times[t], 1, nullptr, all ? nullptr : &active);
      gen_compare_launch c{};
      c.model = sys->model;
      c.pars = sys->d_pars.get<double>();
      c.n_pars = sys->info.n_pars;
      c.n_particles = static_cast<int>(h->P);
      c.n_total = static_cast<int>(n_total);
      c.time = sys->time;
      c.state = sys->d_state.get<double>();
      c.data = h->d_data.get<double>() + t * h->n_groups_data * nd;
      c.shared_data = h->n_groups_data == 1 ? 1 : 0;
      c.out = h->d_ll.get<double>();
      c.accumulate = 1;
      c.active = all ? nullptr : sys->d_active.get<uint8_t>();
      hip_check(launch_general_compare(c, s), "compare kernel");
      if (save_trajectories) {
        hip_check(hipMemcpyAsync(h->d_traj.get<double>() + t * block, sys->d_state.get(),
                                 block * sizeof(double), hipMemcpyDeviceToDevice, s),
                  "copy trajectory");
      }
    }
    std::vector<double> ll(n_total);
    to_host(ll.data(), h->d_ll, n_total * sizeof(double), s);
    for (size_t i = 0; i < groups.size(); ++i) {
      std::copy(ll.data() + groups[i] * h->P, ll.data() + (groups[i] + 1) * h->P,
                ll_out + i * h->P);
    }
    h->last_groups = groups;
    h->saved = std::move(req);
    h->has_run = true;
    return 0;
  } catch (const std::exception& e) {
    return fail(e, err, errlen);
  }
}

int dust2_unfilter_ode_update_pars(dust2_unfilter_ode* h, const double* pars, char* err,
                                   size_t errlen) {
  try {
    check_handle(h, TAG_UNFILTER_ODE, "unfilter");
    return dust2_system_ode_update_pars(h->sys, pars, err, errlen);
  } catch (const std::exception& e) {
    return fail(e, err, errlen);
  }
}

int dust2_unfilter_ode_last_state(dust2_unfilter_ode* h, double* out, char* err,
                                  size_t errlen) {
  try {
    check_handle(h, TAG_UNFILTER_ODE, "unfilter");
    if (!h->has_run) throw std::runtime_error("System has not been run");
    h->sys->errors_pending = false;
    return dust2_system_ode_state(h->sys, out, err, errlen);
  } catch (const std::exception& e) {
    return fail(e, err, errlen);
  }
}

// the states of index_state at every data time of the groups of the last
// run: out[n_time][n_groups_run][P][states]
int dust2_unfilter_ode_last_trajectories(dust2_unfilter_ode* h, double* out, char* err,
                                         size_t errlen) {
  try {
    check_handle(h, TAG_UNFILTER_ODE, "unfilter");
    if (!h->saved.traj) throw std::runtime_error("Trajectories are not current");
    const size_t n_total = h->G * h->P, S = h->sys->S;
    std::vector<double> tmp(h->n_time * S * n_total);
    to_host(tmp.data(), h->d_traj, tmp.size() * sizeof(double), h->sys->stream);
    export_saved_states(tmp.data(), h->n_time, S, n_total, h->P, h->last_groups, nullptr,
                        h->saved.states, out);
    return 0;
  } catch (const std::exception& e) {
    return fail(e, err, errlen);
  }
}

// the whole state at the snapshot times of the last run:
// out[n_snapshots][n_groups_run][P][S]
int dust2_unfilter_ode_last_snapshots(dust2_unfilter_ode* h, double* out, char* err,
                                      size_t errlen) {
  try {
    check_handle(h, TAG_UNFILTER_ODE, "unfilter");
    if (h->saved.n_snap == 0) throw std::runtime_error("Snapshots are not current");
    const size_t n_total = h->G * h->P, S = h->sys->S;
    std::vector<double> tmp(h->n_time * S * n_total);
    to_host(tmp.data(), h->d_traj, tmp.size() * sizeof(double), h->sys->stream);
    export_saved_states(tmp.data(), h->n_time, S, n_total, h->P, h->last_groups,
                        &h->saved.slot, {}, out);
    return 0;
  } catch (const std::exception& e) {
    return fail(e, err, errlen);
  }
}

void dust2_unfilter_ode_free(dust2_unfilter_ode* h) {
  if (handle_ok(h, TAG_UNFILTER_ODE)) delete h;
}

}  // extern "C"

// ============================================================================
// Deterministic discrete-time SIR unfilter with the adjoint gradient
// ============================================================================
struct dust2_unfilter_sir {
  uint32_t tag = TAG_UNFILTER_SIR;
  size_t G = 0, P = 0, n_time = 0, n_groups_data = 0, total_steps = 0;
  double time_start = 0, dt = 1;
  std::vector<double> pars;           // G x 5 (host rows)
  std::vector<double> times;
  std::vector<uint8_t> grad_current;  // per group: gradient of the last run is valid
  std::vector<size_t> last_index;     // groups of the last run
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double last_ms = 0;
  dbuf d_pars, d_sched_time, d_sched_steps, d_obs, d_lfact, d_state, d_ll, d_errors,
      d_counters, d_active, d_hist, d_inc, d_grad, d_traj;
  saved_run saved;                    // trajectories and snapshots of the last run
  bool exact_gradient = false;        // the reference's adjoint unless asked
  ~dust2_unfilter_sir() {
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace {
void check_unfilter_sir(dust2_unfilter_sir* h) {
  check_handle(h, TAG_UNFILTER_SIR, "unfilter");
}

// N, I0, beta, gamma, exp_noise -> N, I0, beta, gamma, exp_noise,
// 1 - exp(-gamma dt), exp(-gamma dt), exp(-gamma)
void sir_adjoint_pars(const double* row, double dt, double* out) {
  for (int i = 0; i < 5; ++i) out[i] = row[i];
  out[5] = 1 - std::exp(-row[3] * dt);
  out[6] = std::exp(-row[3] * dt);
  out[7] = std::exp(-row[3]);
}

void upload_sir_adjoint_pars(dust2_unfilter_sir* h) {
  std::vector<double> kp(h->G * 8);
  for (size_t g = 0; g < h->G; ++g) sir_adjoint_pars(&h->pars[g * kSirPars], h->dt, &kp[g * 8]);
  to_device(h->d_pars, kp.data(), kp.size() * sizeof(double), h->stream);
}
}  // namespace

extern "C" {

int dust2_unfilter_sir_alloc(const double* pars, size_t n_groups, size_t n_particles,
                             double time_start, const double* time, size_t n_time, double dt,
                             const double* data, size_t n_groups_data,
                             dust2_unfilter_sir** out, char* err, size_t errlen) {
  try {
    check_sizes(n_groups, n_particles);
    check_dt(dt);
    check_time_sequence(time_start, time, n_time);
    if (n_groups_data != 1 && n_groups_data != n_groups) {
      throw std::runtime_error("Expected 'data' to have 1 or " + std::to_string(n_groups) +
                               " groups, but it had " + std::to_string(n_groups_data));
    }
    if (n_groups * n_particles > (size_t(1) << 31)) {
      throw std::runtime_error("too many particles");
    }
    for (size_t g = 0; g < n_groups; ++g) check_sir_pars(pars + g * kSirPars);
    auto h = std::make_unique<dust2_unfilter_sir>();
    h->G = n_groups;
    h->P = n_particles;
    h->n_time = n_time;
    h->n_groups_data = n_groups_data;
    h->time_start = time_start;
    h->dt = dt;
    h->pars.assign(pars, pars + n_groups * kSirPars);
    h->times.assign(time, time + n_time);
    h->grad_current.assign(n_groups, 0);
    // interval schedule of the system's run_to_time calls
    // (dust2/discrete/system.hpp:61-92): steps = round(max(0, t - time) / dt),
    // time advanced by steps * dt
    std::vector<double> st(n_time);
    std::vector<uint32_t> ns(n_time);
    double tcur = time_start;
    for (size_t t = 0; t < n_time; ++t) {
      const double n = std::round(std::max(0.0, time[t] - tcur) / dt);
      st[t] = tcur;
      ns[t] = static_cast<uint32_t>(n);
      tcur = tcur + n * dt;
      h->total_steps += ns[t];
    }
    hip_check(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking), "hipStreamCreate");
    hip_check(hipEventCreate(&h->ev0), "hipEventCreate");
    hip_check(hipEventCreate(&h->ev1), "hipEventCreate");
    const size_t n_total = n_groups * n_particles;
    const size_t nd = n_time * n_groups_data;
    std::vector<double> lfact(nd);
    for (size_t i = 0; i < nd; ++i) lfact[i] = std::isnan(data[i]) ? 0.0 : std::lgamma(data[i] + 1);
    h->d_pars.alloc(n_groups * 8 * sizeof(double));
    h->d_sched_time.alloc(n_time * sizeof(double));
    h->d_sched_steps.alloc(n_time * sizeof(uint32_t));
    h->d_obs.alloc(nd * sizeof(double));
    h->d_lfact.alloc(nd * sizeof(double));
    h->d_state.alloc(n_total * 5 * sizeof(double));
    h->d_ll.alloc(n_total * sizeof(double));
    h->d_errors.alloc(n_total * sizeof(error_record));
    h->d_counters.alloc(4 * sizeof(uint32_t));
    h->d_active.alloc(n_groups);
    upload_sir_adjoint_pars(h.get());
    to_device(h->d_sched_time, st.data(), n_time * sizeof(double), h->stream);
    to_device(h->d_sched_steps, ns.data(), n_time * sizeof(uint32_t), h->stream);
    to_device(h->d_obs, data, nd * sizeof(double), h->stream);
    to_device(h->d_lfact, lfact.data(), nd * sizeof(double), h->stream);
    hip_check(hipMemsetAsync(h->d_errors.get(), 0, h->d_errors.bytes(), h->stream), "memset");
    hip_check(hipMemsetAsync(h->d_ll.get(), 0, h->d_ll.bytes(), h->stream), "memset");
    hip_check(hipMemsetAsync(h->d_state.get(), 0, h->d_state.bytes(), h->stream), "memset");
    hip_check(hipStreamSynchronize(h->stream), "sync");
    *out = h.release();
    return 0;
  } catch (const std::exception& e) {
    return fail(e, err, errlen);
  }
}

int dust2_unfilter_sir_run(dust2_unfilter_sir* h, const double* initial, int save_trajectories,
                           const double* save_snapshots, size_t n_snapshots, int adjoint,
                           const size_t* index_state, size_t n_index_state,
                           const size_t* index_group, size_t n_index, double* ll_out,
                           char* err, size_t errlen) {
  try {
    check_unfilter_sir(h);
    const auto idx = check_index(index_group, n_index, h->G);
    saved_run req = check_save(save_trajectories, save_snapshots, n_snapshots, index_state,
                               n_index_state, 5, h->times);
    const bool save = req.any();
    hipStream_t s = h->stream;
    const size_t n_total = h->G * h->P;
    std::vector<uint8_t> active(h->G, 0);
    for (size_t g : idx) active[g] = 1;
    to_device(h->d_active, active.data(), h->G, s);
    if (initial != nullptr) {
      // [G][P][S] -> [S][G*P]
      std::vector<double> tmp(n_total * 5);
      for (size_t k = 0; k < n_total; ++k) {
        for (size_t c = 0; c < 5; ++c) tmp[c * n_total + k] = initial[k * 5 + c];
      }
      to_device(h->d_state, tmp.data(), tmp.size() * sizeof(double), s);
    }
    if (adjoint) {
      // history of S and I at every step and incidence at every data time
      // (dust2/adjoint_data.hpp:28-46), allocated on first use
      const size_t hist_bytes = std::max<size_t>(h->total_steps, 1) * 2 * n_total * sizeof(double);
      if (h->d_hist.bytes() != hist_bytes) {
        h->d_hist.alloc(hist_bytes);
        h->d_inc.alloc(h->n_time * n_total * sizeof(double));
        h->d_grad.alloc(n_total * 3 * sizeof(double));
      }
    }
    if (save && h->d_traj.bytes() == 0) {
      h->d_traj.alloc(h->n_time * 5 * n_total * sizeof(double));
    }
    hip_check(hipMemsetAsync(h->d_counters.get(), 0, h->d_counters.bytes(), s), "memset");
    sir_unfilter_launch a{};
    a.traj = save ? h->d_traj.get<double>() : nullptr;
    a.pars = h->d_pars.get<double>();
    a.active = h->d_active.get<uint8_t>();
    a.n_particles = static_cast<int>(h->P);
    a.n_total = static_cast<int>(n_total);
    a.set_initial = initial == nullptr ? 1 : 0;
    a.dt = h->dt;
    a.n_time = static_cast<int>(h->n_time);
    a.sched_time = h->d_sched_time.get<double>();
    a.sched_steps = h->d_sched_steps.get<uint32_t>();
    a.obs = h->d_obs.get<double>();
    a.lfact = h->d_lfact.get<double>();
    a.obs_stride = h->n_groups_data;
    a.obs_group_stride = h->n_groups_data == 1 ? 0 : 1;
    a.state = h->d_state.get<double>();
    a.ll = h->d_ll.get<double>();
    a.hist = adjoint ? h->d_hist.get<double>() : nullptr;
    a.inc = adjoint ? h->d_inc.get<double>() : nullptr;
    a.grad = adjoint ? h->d_grad.get<double>() : nullptr;
    a.errors = h->d_errors.get<error_record>();
    a.error_count = h->d_counters.get<uint32_t>();
    a.exact = h->exact_gradient ? 1 : 0;
    hip_check(hipEventRecord(h->ev0, s), "hipEventRecord");
    hip_check(launch_sir_unfilter(a, s), "sir unfilter kernel");
    hip_check(hipEventRecord(h->ev1, s), "hipEventRecord");
    uint32_t count = 0;
    to_host(&count, h->d_counters, sizeof(count), s);
    float ms = 0;
    hip_check(hipEventElapsedTime(&ms, h->ev0, h->ev1), "hipEventElapsedTime");
    h->last_ms = ms;
    for (size_t g : idx) h->grad_current[g] = 0;
    h->saved = saved_run{};
    if (count > 0) {
      std::vector<error_record> rec(n_total);
      to_host(rec.data(), h->d_errors, rec.size() * sizeof(error_record), s);
      hip_check(hipMemsetAsync(h->d_errors.get(), 0, h->d_errors.bytes(), s), "memset");
      hip_check(hipStreamSynchronize(s), "sync");
      throw std::runtime_error(error_report(rec, count));
    }
    if (adjoint) {
      for (size_t g : idx) h->grad_current[g] = 1;
    }
    h->saved = std::move(req);
    h->last_index = idx;
    std::vector<double> ll(n_total);
    to_host(ll.data(), h->d_ll, ll.size() * sizeof(double), s);
    for (size_t i = 0; i < idx.size(); ++i) {
      std::memcpy(ll_out + i * h->P, ll.data() + idx[i] * h->P, h->P * sizeof(double));
    }
    return 0;
  } catch (const std::exception& e) {
    return fail(e, err, errlen);
  }
}

int dust2_unfilter_sir_last_gradient(dust2_unfilter_sir* h, double* out, char* err,
                                     size_t errlen) {
  try {
    check_unfilter_sir(h);
    bool any = false;
    for (size_t g : h->last_index) any = any || h->grad_current[g];
    if (!any) {
      throw std::runtime_error("System was not run with 'adjoint = TRUE'");
    }
    if (h->P > 1) {
      throw std::runtime_error("n_particles > 1 not yet supported; see mrc-5406");
    }
    std::vector<double> grad(h->G * 3);
    to_host(grad.data(), h->d_grad, grad.size() * sizeof(double), h->stream);
    for (size_t i = 0; i < h->last_index.size(); ++i) {
      std::memcpy(out + 3 * i, grad.data() + 3 * h->last_index[i], 3 * sizeof(double));
    }
    return 0;
  } catch (const std::exception& e) {
    return fail(e, err, errlen);
  }
}

int dust2_unfilter_sir_last_trajectories(dust2_unfilter_sir* h, double* out, char* err,
                                         size_t errlen) {
  try {
    check_unfilter_sir(h);
    if (!h->saved.traj) throw std::runtime_error("Trajectories are not current");
    const size_t NT = h->G * h->P, T = h->n_time;
    std::vector<double> tmp(T * 5 * NT);
    to_host(tmp.data(), h->d_traj, tmp.size() * sizeof(double), h->stream);
    export_saved_states(tmp.data(), T, 5, NT, h->P, h->last_index, nullptr, h->saved.states,
                        out);
    return 0;
  } catch (const std::exception& e) {
    return fail(e, err, errlen);
  }
}

int dust2_unfilter_sir_last_snapshots(dust2_unfilter_sir* h, double* out, char* err,
                                      size_t errlen) {
  try {
    check_unfilter_sir(h);
    if (h->saved.n_snap == 0) throw std::runtime_error("Snapshots are not current");
    const size_t NT = h->G * h->P, T = h->n_time;
    std::vector<double> tmp(T * 5 * NT);
    to_host(tmp.data(), h->d_traj, tmp.size() * sizeof(double), h->stream);
    export_saved_states(tmp.data(), T, 5, NT, h->P, h->last_index, &h->saved.slot, {}, out);
    return 0;
  } catch (const std::exception& e) {
    return fail(e, err, errlen);
  }
}

int dust2_unfilter_sir_update_pars(dust2_unfilter_sir* h, const double* pars,
                                   const size_t* index_group, size_t n_index, char* err,
                                   size_t errlen) {
  try {
    check_unfilter_sir(h);
    const auto idx = check_index(index_group, n_index, h->G);
    for (size_t i = 0; i < idx.size(); ++i) {
      const double* row = pars + i * kSirPars;
      check_sir_pars(row);
      double* dst = &h->pars[idx[i] * kSirPars];
      // I0, beta and gamma update; N and exp_noise are constant
      // (inst/examples/sir.cpp:89-93)
      dst[1] = row[1];
      dst[2] = row[2];
      dst[3] = row[3];
      h->grad_current[idx[i]] = 0;
    }
    upload_sir_adjoint_pars(h);
    hip_check(hipStreamSynchronize(h->stream), "sync");
    return 0;
  } catch (const std::exception& e) {
    return fail(e, err, errlen);
  }
}

int dust2_unfilter_sir_last_state(dust2_unfilter_sir* h, double* out, char* err,
                                  size_t errlen) {
  try {
    check_unfilter_sir(h);
    const size_t n = h->G * h->P;
    std::vector<double> tmp(n * 5);
    to_host(tmp.data(), h->d_state, tmp.size() * sizeof(double), h->stream);
    for (size_t k = 0; k < n; ++k) {
      for (size_t c = 0; c < 5; ++c) out[k * 5 + c] = tmp[c * n + k];
    }
    return 0;
  } catch (const std::exception& e) {
    return fail(e, err, errlen);
  }
}

int dust2_unfilter_sir_set_adjoint_form(dust2_unfilter_sir* h, int exact, char* err,
                                        size_t errlen) {
  try {
    check_unfilter_sir(h);
    if (exact != 0 && exact != 1) throw std::runtime_error("'exact' must be 0 or 1");
    h->exact_gradient = exact != 0;
    std::fill(h->grad_current.begin(), h->grad_current.end(), 0);
    return 0;
  } catch (const std::exception& e) {
    return fail(e, err, errlen);
  }
}

void dust2_unfilter_sir_free(dust2_unfilter_sir* h) {
  if (handle_ok(h, TAG_UNFILTER_SIR)) delete h;
}

}  // extern "C"

// ============================================================================
// Serving: continuous batching of likelihood evaluations
// ============================================================================
namespace {

// First-fit allocator over one block of memory (256-byte granules): the
// server's device HBM for resident requests and its pinned host staging for
// their uploads, so admitting and retiring requests never allocates.
class arena {
 public:
  void init(char* base, size_t bytes) {
    base_ = base;
    cap_ = bytes;
    free_.clear();
    free_[0] = bytes;
    used_ = 0;
  }
  bool alloc(size_t bytes, size_t& off) {
    bytes = (bytes + 255) & ~size_t(255);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second >= bytes) {
        off = it->first;
        const size_t rest = it->second - bytes;
        free_.erase(it);
        if (rest) free_[off + bytes] = rest;
        used_ += bytes;
        return true;
      }
    }
    return false;
  }
  void release(size_t off, size_t bytes) {
    bytes = (bytes + 255) & ~size_t(255);
    used_ -= bytes;
    auto it = free_.emplace(off, bytes).first;
    auto next = std::next(it);
    if (next != free_.end() && it->first + it->second == next->first) {
      it->second += next->second;
      free_.erase(next);
    }
    if (it != free_.begin()) {
      auto prev = std::prev(it);
      if (prev->first + prev->second == it->first) {
        prev->second += it->second;
        free_.erase(it);
      }
    }
  }
  char* base() const { return base_; }
  size_t used() const { return used_; }
  size_t capacity() const { return cap_; }

 private:
  char* base_ = nullptr;
  size_t cap_ = 0;
  size_t used_ = 0;
  std::map<size_t, size_t> free_;
};

struct srv_request {
  uint64_t id = 0;
  int model = 0;  // 0 sir, 1 seir16, 2 sirode
  size_t P = 0;
  double time_start = 0, dt = 1;
  std::vector<double> pars, times, data;
  uint64_t seed = 0;
  int64_t chain = -1;
};

struct srv_result {
  uint64_t id;
  double ll;
  int status;
  std::string message;
};

// Sub-buffer layout of one resident request inside its arena block; the
// region from `upload` to `total` is filled by one copy from pinned staging.
struct srv_layout {
  size_t state_a, state_b, rng, lambda, anc, errors, ll, counters, upload, sched_time,
      sched_steps, sched_zero, obs, lfact, kpars, total;
};

size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

srv_layout srv_plan(size_t S, size_t P, size_t T, size_t n_data, size_t kpars, bool own_rng) {
  srv_layout L{};
  size_t o = 0;
  auto take = [&](size_t bytes) {
    const size_t at = o;
    o += align256(bytes);
    return at;
  };
  L.state_a = take(S * P * 4);
  L.state_b = take(S * P * 4);
  L.rng = own_rng ? take(4 * (P + 2) * 8) : 0;
  L.lambda = take(P * 8);
  L.anc = take(P * 4);
  L.errors = take(P * sizeof(error_record));
  L.ll = take(8);
  L.counters = take(8);
  L.upload = o;
  L.sched_time = take(T * 8);
  L.sched_steps = take(T * 4);
  L.sched_zero = take(T * 8);
  L.obs = take(T * n_data * 8);
  L.lfact = take(T * n_data * 8);
  L.kpars = take(kpars * 8);
  L.total = o;
  return L;
}

struct srv_slot {
  srv_request req;
  size_t off = 0;        // device arena block
  size_t stage_off = 0;  // pinned staging block
  srv_layout L{};
  size_t T = 0;
  size_t done = 0;       // intervals run
  pipe_slot ps{};
};

// An MCMC chain's filter streams, kept resident between its requests so
// that each request continues the chain's streams (a persistent filter per
// chain, as dust_likelihood_monty keeps one per chain, R/monty.R:153-204).
// The streams start at long_jump^chain of the seed (R/monty.R:29-36).
struct srv_chain {
  size_t P = 0;
  size_t off = 0;
  bool seeded = false;
  bool busy = false;
};

// The resident requests of one model, whatever their particle counts and
// data lengths: one stream, one device slot table of fixed capacity
// (unused entries do nothing) with a block map that lays every member's
// particles onto the advance launch's workgroups, a HIP graph of
// chunk x (advance over the block map, resample with one workgroup per
// member) replayed per launch, and the completion event of the launch in
// flight.
struct srv_group {
  int model = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev = nullptr;
  bool in_flight = false;
  std::vector<std::unique_ptr<srv_slot>> members;  // table order at the next launch
  std::vector<srv_slot*> launched;                 // members of the launch in flight
  size_t cap = 0;   // slot table entries
  size_t bcap = 0;  // block map entries
  dbuf d_table, d_status, d_bmap;
  pinned h_table, h_status, h_bmap;
  hipGraphExec_t graph = nullptr;
  int graph_chunk = 0;
  int graph_ppt = 0;  // particles per thread of the captured resample launches
  ~srv_group() {
    if (graph) (void)hipGraphExecDestroy(graph);
    if (ev) (void)hipEventDestroy(ev);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

// Deterministic (ODE) requests: every queued one in one launch, each with
// its own series; inputs packed into pinned staging, results copied back
// into pinned memory, completion seen through an event.
struct srv_ode {
  hipStream_t stream = nullptr;
  hipEvent_t ev = nullptr;
  bool in_flight = false;
  std::vector<srv_request> launched;
  pinned h_in, h_out;
  dbuf d_in, d_out;
  ~srv_ode() {
    if (ev) (void)hipEventDestroy(ev);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

}  // namespace

struct dust2_server {
  uint32_t tag = TAG_SERVER;
  dbuf d_mem;
  pinned h_mem;
  arena mem;    // device
  arena stage;  // pinned host
  int chunk = 8;
  uint64_t next_id = 1;
  std::deque<srv_request> queue;
  std::deque<srv_request> ode_queue;
  std::map<int, std::unique_ptr<srv_group>> groups;  // by model
  std::map<int64_t, srv_chain> chains;
  srv_ode ode;
  std::vector<srv_result> results;
  size_t max_ode_per_launch = 65536;
};

namespace {

// Server model numbers: 0 sir, 1 seir16 (the discrete-time model ids of
// models.hpp), 2 sirode.
constexpr int kServerOde = 2;
size_t model_state(int model) { return discrete_model(model).n_state(nullptr); }
size_t model_n_pars(int model) {
  return model == kServerOde ? kSirodePars : discrete_model(model).n_pars;
}
size_t model_n_data(int model) { return model == kServerOde ? 1 : discrete_model(model).n_data; }
size_t model_kpars(int model) { return discrete_model(model).kpars; }
double model_zero_period(int model) { return discrete_model(model).zero_period; }

srv_group& srv_group_of(dust2_server* sv, int model) {
  auto& g = sv->groups[model];
  if (!g) {
    g = std::make_unique<srv_group>();
    g->model = model;
    hip_check(hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking), "hipStreamCreate");
    hip_check(hipEventCreateWithFlags(&g->ev, hipEventDisableTiming), "hipEventCreate");
  }
  return *g;
}

// Place a request in the arena and queue the uploads and seeding of its
// resident buffers on its group's stream: no allocation and no wait.
bool srv_admit(dust2_server* sv, const srv_request& rq) {
  const size_t S = model_state(rq.model), P = rq.P, T = rq.times.size();
  const size_t nd = model_n_data(rq.model), kp = model_kpars(rq.model);
  const bool own_rng = rq.chain < 0;
  const srv_layout L = srv_plan(S, P, T, nd, kp, own_rng);
  // a chain keeps one block of streams for its particle count
  srv_chain* ch = nullptr;
  if (!own_rng) {
    ch = &sv->chains[rq.chain];
    if (ch->P != P) {
      if (ch->P != 0) sv->mem.release(ch->off, 4 * (ch->P + 2) * 8);
      ch->P = 0;
      ch->seeded = false;
      if (!sv->mem.alloc(4 * (P + 2) * 8, ch->off)) return false;
      ch->P = P;
    }
  }
  size_t off = 0, soff = 0;
  if (!sv->mem.alloc(L.total, off)) return false;
  const size_t up = L.total - L.upload;
  if (!sv->stage.alloc(up, soff)) {
    sv->mem.release(off, L.total);
    return false;
  }
  auto slot = std::make_unique<srv_slot>();
  slot->req = rq;
  slot->off = off;
  slot->stage_off = soff;
  slot->L = L;
  slot->T = T;
  char* b = sv->mem.base() + off;
  char* h = sv->stage.base() + soff;
  // schedule (dust2/discrete/system.hpp:61-92), data and parameters into
  // the staging block, laid out as the device block from `upload`
  double* st = reinterpret_cast<double*>(h + (L.sched_time - L.upload));
  uint32_t* ns = reinterpret_cast<uint32_t*>(h + (L.sched_steps - L.upload));
  uint64_t* zm = reinterpret_cast<uint64_t*>(h + (L.sched_zero - L.upload));
  double* ob = reinterpret_cast<double*>(h + (L.obs - L.upload));
  double* lf = reinterpret_cast<double*>(h + (L.lfact - L.upload));
  double* kpars = reinterpret_cast<double*>(h + (L.kpars - L.upload));
  double tcur = rq.time_start;
  const double period = model_zero_period(rq.model);
  for (size_t t = 0; t < T; ++t) {
    const size_t n = static_cast<size_t>(std::round(std::max(0.0, rq.times[t] - tcur) / rq.dt));
    st[t] = tcur;
    ns[t] = static_cast<uint32_t>(n);
    uint64_t m = 0;
    for (size_t k = 0; k < std::min<size_t>(n, 64); ++k) {
      if (std::fmod(tcur + k * rq.dt, period) == 0) m |= uint64_t(1) << k;
    }
    zm[t] = m;
    tcur = tcur + n * rq.dt;
  }
  for (size_t i = 0; i < T * nd; ++i) {
    ob[i] = rq.data[i];
    lf[i] = std::isnan(rq.data[i]) ? 0.0 : std::lgamma(rq.data[i] + 1);
  }
  std::fill(kpars, kpars + kp, 0.0);
  discrete_model(rq.model).kernel_pars(rq.pars.data(), rq.dt, kpars);
  srv_group& g = srv_group_of(sv, rq.model);
  hipStream_t s = g.stream;
  hip_check(hipMemcpyAsync(b + L.upload, h, up, hipMemcpyHostToDevice, s), "upload");
  hip_check(hipMemsetAsync(b + L.ll, 0, 8, s), "memset");
  hip_check(hipMemsetAsync(b + L.counters, 0, 8, s), "memset");
  hip_check(hipMemsetAsync(b + L.errors, 0, P * sizeof(error_record), s), "memset");
  uint64_t* rng = own_rng ? reinterpret_cast<uint64_t*>(b + L.rng)
                          : reinterpret_cast<uint64_t*>(sv->mem.base() + ch->off);
  if (own_rng || !ch->seeded) {
    host_rng r;
    host_seed_words(rq.seed, r.s);
    for (int64_t c = 0; !own_rng && c < rq.chain; ++c) host_long_jump(r);
    hip_check(launch_seed_streams(r.s, tables_for_current_device().jump.get<uint64_t>(),
                                  kJumpMats, P + 2, rng, P + 2, s),
              "seed streams");
  }
  if (ch) {
    ch->seeded = true;
    ch->busy = true;
  }
  pipe_slot& ps = slot->ps;
  ps.pars = reinterpret_cast<const double*>(b + L.kpars);
  ps.obs = reinterpret_cast<const double*>(b + L.obs);
  ps.lfact = reinterpret_cast<const double*>(b + L.lfact);
  ps.obs_stride = static_cast<uint32_t>(nd);
  ps.n_particles = static_cast<uint32_t>(P);
  ps.flags = SLOT_AUTO_MODE;
  ps.group = 0;
  ps.sched_time = reinterpret_cast<const double*>(b + L.sched_time);
  ps.sched_steps = reinterpret_cast<const uint32_t*>(b + L.sched_steps);
  ps.sched_zero = reinterpret_cast<const uint64_t*>(b + L.sched_zero);
  ps.dt = rq.dt;
  ps.state_a = b + L.state_a;
  ps.state_b = b + L.state_b;
  ps.rng = rng;
  ps.lambda = reinterpret_cast<double*>(b + L.lambda);
  ps.ancestors = reinterpret_cast<uint32_t*>(b + L.anc);
  ps.ll = reinterpret_cast<double*>(b + L.ll);
  ps.errors = reinterpret_cast<error_record*>(b + L.errors);
  ps.error_count = reinterpret_cast<uint32_t*>(b + L.counters);
  ps.n_fallback = reinterpret_cast<uint32_t*>(b + L.counters) + 1;
  ps.t_offset = 0;
  ps.n_time = static_cast<uint32_t>(T);
  g.members.push_back(std::move(slot));
  return true;
}

// Launch the group's next chunk of intervals for every member.
void srv_launch(dust2_server* sv, srv_group& g) {
  const size_t n = g.members.size();
  if (n >= (size_t(1) << (32 - kBlockBits))) {
    throw std::runtime_error("too many resident requests for one launch");
  }
  bool recapture = !g.graph || g.graph_chunk != sv->chunk;
  if (n > g.cap) {
    size_t cap = std::max<size_t>(16, g.cap);
    while (cap < n) cap *= 2;
    g.d_table.alloc(cap * sizeof(pipe_slot));
    g.h_table.alloc(cap * sizeof(pipe_slot));
    g.d_status.alloc(2 * cap * sizeof(double));
    g.h_status.alloc(2 * cap * sizeof(double));
    g.cap = cap;
    recapture = true;
  }
  // the members' particles in runs of the advance workgroup's size
  const size_t B = static_cast<size_t>(discrete_advance_block(g.model, false));
  size_t n_blocks = 0, max_p = 1;
  for (const auto& sl : g.members) {
    n_blocks += (sl->req.P + B - 1) / B;
    max_p = std::max(max_p, sl->req.P);
  }
  if (n_blocks > g.bcap) {
    size_t bcap = std::max<size_t>(256, g.bcap);
    while (bcap < n_blocks) bcap *= 2;
    g.d_bmap.alloc(bcap * sizeof(uint32_t));
    g.h_bmap.alloc(bcap * sizeof(uint32_t));
    g.bcap = bcap;
    recapture = true;
  }
  // the resample workgroup holds the largest member; recaptured when a
  // larger one arrives
  const int ppt = pipe_resample_ppt(static_cast<int>(max_p));
  if (ppt > g.graph_ppt) recapture = true;
  const int cap = static_cast<int>(g.cap);
  if (recapture) {
    if (g.graph) hip_check(hipGraphExecDestroy(g.graph), "hipGraphExecDestroy");
    g.graph = nullptr;
    g.graph_ppt = std::max(g.graph_ppt, ppt);
    const int p_res = 1024 * g.graph_ppt;
    hipGraph_t graph;
    hip_check(hipStreamBeginCapture(g.stream, hipStreamCaptureModeThreadLocal), "begin capture");
    for (int k = 0; k < sv->chunk; ++k) {
      hip_check(launch_discrete_advance_mapped(g.model, false, g.d_table.get<pipe_slot>(),
                                               g.d_bmap.get<uint32_t>(),
                                               static_cast<int>(g.bcap), k, g.stream),
                "advance kernel");
      hip_check(launch_pipe_resample(g.d_table.get<pipe_slot>(), cap, p_res, k, g.stream),
                "resample kernel");
    }
    hip_check(hipStreamEndCapture(g.stream, &graph), "end capture");
    hip_check(hipGraphInstantiate(&g.graph, graph, nullptr, nullptr, 0), "graph instantiate");
    hip_check(hipGraphDestroy(graph), "hipGraphDestroy");
    g.graph_chunk = sv->chunk;
  }
  pipe_slot* tab = g.h_table.get<pipe_slot>();
  uint32_t* bmap = g.h_bmap.get<uint32_t>();
  size_t nb = 0;
  g.launched.clear();
  for (size_t i = 0; i < n; ++i) {
    srv_slot* sl = g.members[i].get();
    tab[i] = sl->ps;
    tab[i].t_offset = static_cast<int32_t>(sl->done);
    g.launched.push_back(sl);
    const size_t blocks = (sl->req.P + B - 1) / B;
    for (size_t b = 0; b < blocks; ++b) {
      bmap[nb++] = static_cast<uint32_t>(i << kBlockBits | b);
    }
  }
  for (size_t i = n; i < g.cap; ++i) {
    // unused entries: no particles and no intervals, so every kernel leaves
    tab[i] = pipe_slot{};
    tab[i].flags = SLOT_AUTO_MODE;
  }
  std::fill(bmap + nb, bmap + g.bcap, kNoBlock);
  hip_check(hipMemcpyAsync(g.d_table.get(), tab, g.cap * sizeof(pipe_slot),
                           hipMemcpyHostToDevice, g.stream),
            "upload slot table");
  hip_check(hipMemcpyAsync(g.d_bmap.get(), bmap, g.bcap * sizeof(uint32_t),
                           hipMemcpyHostToDevice, g.stream),
            "upload block map");
  hip_check(hipGraphLaunch(g.graph, g.stream), "hipGraphLaunch");
  hip_check(launch_slot_status(g.d_table.get<pipe_slot>(), static_cast<int>(n),
                               g.d_status.get<double>(), g.stream),
            "status kernel");
  hip_check(hipMemcpyAsync(g.h_status.get(), g.d_status.get(), 2 * n * sizeof(double),
                           hipMemcpyDeviceToHost, g.stream),
            "status copy");
  hip_check(hipEventRecord(g.ev, g.stream), "hipEventRecord");
  g.in_flight = true;
}

// Results of a completed launch: finished requests leave their slots.
void srv_complete(dust2_server* sv, srv_group& g) {
  const double* status = g.h_status.get<double>();
  std::vector<srv_slot*> finished;
  for (size_t i = 0; i < g.launched.size(); ++i) {
    srv_slot* sl = g.launched[i];
    sl->done += static_cast<size_t>(sv->chunk);
    const uint32_t count = static_cast<uint32_t>(status[2 * i + 1]);
    if (sl->done < sl->T && count == 0) continue;
    srv_result r{sl->req.id, status[2 * i], 0, std::string()};
    if (count > 0) {
      std::vector<error_record> rec(sl->req.P);
      hip_check(hipMemcpy(rec.data(), sv->mem.base() + sl->off + sl->L.errors,
                          rec.size() * sizeof(error_record), hipMemcpyDeviceToHost),
                "copy");
      r.status = 1;
      r.message = error_report(rec, count);
    }
    sv->results.push_back(std::move(r));
    finished.push_back(sl);
  }
  g.launched.clear();
  g.in_flight = false;
  if (finished.empty()) return;
  std::vector<std::unique_ptr<srv_slot>> keep;
  for (auto& sl : g.members) {
    if (std::find(finished.begin(), finished.end(), sl.get()) == finished.end()) {
      keep.push_back(std::move(sl));
      continue;
    }
    sv->mem.release(sl->off, sl->L.total);
    sv->stage.release(sl->stage_off, sl->L.total - sl->L.upload);
    if (sl->req.chain >= 0) sv->chains[sl->req.chain].busy = false;
  }
  g.members.swap(keep);
}

// All queued deterministic (ODE) requests in one launch, each with its own
// series and start time (one thread per request).
void srv_launch_ode(dust2_server* sv) {
  srv_ode& o = sv->ode;
  const size_t n = std::min(sv->ode_queue.size(), sv->max_ode_per_launch);
  size_t n_t = 0;
  for (size_t i = 0; i < n; ++i) n_t += sv->ode_queue[i].times.size();
  // packed inputs: pars [n][4], t0 [n], off [n], len [n], times, obs, lfact
  const size_t b_pars = align256(n * kSirodePars * 8), b_t0 = align256(n * 8),
               b_off = align256(n * 8), b_len = align256(n * 4), b_t = align256(n_t * 8);
  const size_t in_bytes = b_pars + b_t0 + b_off + b_len + 3 * b_t;
  // results: ll [n], errors [n], state [5][n], counters
  const size_t b_ll = align256(n * 8), b_err = align256(n * sizeof(error_record)),
               b_state = align256(n * 5 * 8);
  const size_t out_bytes = b_ll + b_err + b_state + 256;
  if (o.h_in.bytes() < in_bytes) o.h_in.alloc(std::max(in_bytes, 2 * o.h_in.bytes()));
  if (o.d_in.bytes() < in_bytes) o.d_in.alloc(o.h_in.bytes());
  if (o.h_out.bytes() < out_bytes) o.h_out.alloc(std::max(out_bytes, 2 * o.h_out.bytes()));
  if (o.d_out.bytes() < out_bytes) o.d_out.alloc(o.h_out.bytes());
  char* h = o.h_in.get<char>();
  double* pars = reinterpret_cast<double*>(h);
  double* t0 = reinterpret_cast<double*>(h + b_pars);
  int64_t* off = reinterpret_cast<int64_t*>(h + b_pars + b_t0);
  int32_t* len = reinterpret_cast<int32_t*>(h + b_pars + b_t0 + b_off);
  double* times = reinterpret_cast<double*>(h + b_pars + b_t0 + b_off + b_len);
  double* obs = times + b_t / 8;
  double* lf = obs + b_t / 8;
  o.launched.clear();
  size_t at = 0;
  for (size_t i = 0; i < n; ++i) {
    srv_request& rq = sv->ode_queue.front();
    std::copy(rq.pars.begin(), rq.pars.end(), pars + i * kSirodePars);
    t0[i] = rq.time_start;
    off[i] = static_cast<int64_t>(at);
    len[i] = static_cast<int32_t>(rq.times.size());
    for (size_t k = 0; k < rq.times.size(); ++k, ++at) {
      times[at] = rq.times[k];
      obs[at] = rq.data[k];
      lf[at] = std::isnan(rq.data[k]) ? 0.0 : std::lgamma(rq.data[k] + 1);
    }
    o.launched.push_back(std::move(rq));
    sv->ode_queue.pop_front();
  }
  char* d = o.d_in.get<char>();
  char* dout = o.d_out.get<char>();
  hipStream_t s = o.stream;
  hip_check(hipMemcpyAsync(d, h, in_bytes, hipMemcpyHostToDevice, s), "upload");
  hip_check(hipMemsetAsync(dout + b_ll, 0, b_err, s), "memset");
  hip_check(hipMemsetAsync(dout + b_ll + b_err + b_state, 0, 256, s), "memset");
  ode_launch a{};
  a.sens_param = -1;
  a.pars = reinterpret_cast<const double*>(d);
  a.n_particles = 1;
  a.n_total = static_cast<int>(n);
  a.time_start = 0;
  a.series_time_start = reinterpret_cast<const double*>(d + b_pars);
  a.series_offset = reinterpret_cast<const int64_t*>(d + b_pars + b_t0);
  a.series_len = reinterpret_cast<const int32_t*>(d + b_pars + b_t0 + b_off);
  a.times = reinterpret_cast<const double*>(d + b_pars + b_t0 + b_off + b_len);
  a.obs = a.times + b_t / 8;
  a.lfact = a.obs + b_t / 8;
  a.control = ode_control_dev{10000, 1e-6, 1e-6, std::numeric_limits<double>::epsilon(),
                              std::numeric_limits<double>::infinity(), 0.9, 0.2, 10.0, 0.04,
                              0.2 - 0.04 * 0.75};
  a.ll = reinterpret_cast<double*>(dout);
  a.errors = reinterpret_cast<error_record*>(dout + b_ll);
  a.state = dout + b_ll + b_err;
  a.error_flag = reinterpret_cast<uint32_t*>(dout + b_ll + b_err + b_state);
  hip_check(launch_dp5_unfilter(a, true, s), "dp5 serving kernel");
  hip_check(hipMemcpyAsync(o.h_out.get(), dout, b_ll + b_err, hipMemcpyDeviceToHost, s),
            "result copy");
  hip_check(hipEventRecord(o.ev, s), "hipEventRecord");
  o.in_flight = true;
}

void srv_complete_ode(dust2_server* sv) {
  srv_ode& o = sv->ode;
  const size_t n = o.launched.size();
  const double* ll = o.h_out.get<double>();
  const error_record* rec =
      reinterpret_cast<const error_record*>(o.h_out.get<char>() + align256(n * 8));
  for (size_t i = 0; i < n; ++i) {
    sv->results.push_back(srv_result{o.launched[i].id, ll[i], rec[i].code ? 1 : 0,
                                     rec[i].code ? error_text(rec[i]) : std::string()});
  }
  o.launched.clear();
  o.in_flight = false;
}

// true when the launch behind `ev` has completed
bool event_done(hipEvent_t ev) {
  const hipError_t e = hipEventQuery(ev);
  if (e == hipSuccess) return true;
  if (e == hipErrorNotReady) return false;
  hip_check(e, "hipEventQuery");
  return false;
}

}  // namespace

extern "C" {

int dust2_amd_rng_jump(uint64_t* state, size_t n_streams, int long_jump, char* err,
                       size_t errlen) {
  try {
    if (state == nullptr && n_streams > 0) throw std::runtime_error("'state' is NULL");
    for (size_t i = 0; i < n_streams; ++i) {
      host_rng r;
      std::memcpy(r.s, state + 4 * i, 32);
      if (long_jump) {
        host_long_jump(r);
      } else {
        host_jump(r);
      }
      std::memcpy(state + 4 * i, r.s, 32);
    }
    return 0;
  } catch (const std::exception& e) {
    return fail(e, err, errlen);
  }
}

int dust2_server_create(size_t hbm_bytes, int chunk_intervals, dust2_server** out, char* err,
                        size_t errlen) {
  try {
    if (chunk_intervals < 1) throw std::runtime_error("'chunk_intervals' must be at least 1");
    if (hbm_bytes < (size_t(1) << 20)) throw std::runtime_error("'hbm_bytes' is too small");
    auto sv = std::make_unique<dust2_server>();
    sv->chunk = chunk_intervals;
    sv->d_mem.alloc(hbm_bytes);
    sv->mem.init(sv->d_mem.get<char>(), hbm_bytes);
    // uploads are a small fraction of a request's device footprint
    const size_t stage_bytes = std::max<size_t>(size_t(16) << 20, hbm_bytes / 8);
    sv->h_mem.alloc(stage_bytes);
    sv->stage.init(sv->h_mem.get<char>(), stage_bytes);
    hip_check(hipStreamCreateWithFlags(&sv->ode.stream, hipStreamNonBlocking),
              "hipStreamCreate");
    hip_check(hipEventCreateWithFlags(&sv->ode.ev, hipEventDisableTiming), "hipEventCreate");
    tables_for_current_device();
    *out = sv.release();
    return 0;
  } catch (const std::exception& e) {
    return fail(e, err, errlen);
  }
}

int dust2_server_submit(dust2_server* sv, int model, const double* pars, size_t n_particles,
                        double time_start, const double* time, size_t n_time, double dt,
                        const double* data, uint64_t seed, int64_t chain, uint64_t* id_out,
                        char* err, size_t errlen) {
  try {
    check_handle(sv, TAG_SERVER, "server");
    if (model < 0 || model > 2) throw std::runtime_error("unknown model");
    check_time_sequence(time_start, time, n_time);
    srv_request rq;
    rq.id = sv->next_id++;
    rq.model = model;
    rq.P = model == kServerOde ? 1 : n_particles;
    rq.time_start = time_start;
    rq.dt = dt;
    rq.seed = seed;
    rq.chain = model == kServerOde ? -1 : chain;
    const size_t np = model_n_pars(model);
    rq.pars.assign(pars, pars + np);
    rq.times.assign(time, time + n_time);
    rq.data.assign(data, data + n_time * model_n_data(model));
    if (model == kServerOde) {
      check_sirode_pars(pars);
      sv->ode_queue.push_back(std::move(rq));
    } else {
      check_sizes(1, n_particles);
      check_dt(dt);
      if (pipe_resample_ppt(static_cast<int>(n_particles)) == 0) {
        throw std::runtime_error("'n_particles' too large");
      }
      if (chain > (int64_t(1) << 20)) throw std::runtime_error("'chain' is too large");
      discrete_model(model).check(pars);
      sv->queue.push_back(std::move(rq));
    }
    *id_out = sv->next_id - 1;
    return 0;
  } catch (const std::exception& e) {
    return fail(e, err, errlen);
  }
}

// One scheduling pass: collect the launches that have completed (their
// finished requests become results and free their memory), admit queued
// requests while their footprint fits, and launch the next chunk of every
// idle group and the queued ODE requests.  Groups progress independently,
// each on its own stream; the call waits only when nothing else could be
// done while work is in flight.
int dust2_server_step(dust2_server* sv, size_t* n_active, size_t* n_queued, char* err,
                      size_t errlen) {
  try {
    check_handle(sv, TAG_SERVER, "server");
    for (int pass = 0; pass < 2; ++pass) {
      bool progress = false;
      for (auto& kv : sv->groups) {
        srv_group& g = *kv.second;
        if (g.in_flight && event_done(g.ev)) {
          srv_complete(sv, g);
          progress = true;
        }
      }
      if (sv->ode.in_flight && event_done(sv->ode.ev)) {
        srv_complete_ode(sv);
        progress = true;
      }
      // admission in order of arrival; a request of a chain with a request
      // resident waits for it; a request that does not fit stops admission
      bool resident = false;
      for (auto& kv : sv->groups) resident = resident || !kv.second->members.empty();
      for (auto it = sv->queue.begin(); it != sv->queue.end();) {
        if (it->chain >= 0) {
          auto c = sv->chains.find(it->chain);
          if (c != sv->chains.end() && c->second.busy) {
            ++it;
            continue;
          }
        }
        if (!srv_admit(sv, *it)) {
          if (!resident) {
            throw std::runtime_error("request does not fit the server's memory budget");
          }
          break;
        }
        resident = true;
        progress = true;
        it = sv->queue.erase(it);
      }
      for (auto& kv : sv->groups) {
        srv_group& g = *kv.second;
        if (!g.in_flight && !g.members.empty()) {
          srv_launch(sv, g);
          progress = true;
        }
      }
      if (!sv->ode.in_flight && !sv->ode_queue.empty()) {
        srv_launch_ode(sv);
        progress = true;
      }
      if (progress || pass == 1) break;
      // nothing to do: wait until any launch in flight completes (groups
      // finish in any order, so no single event is waited on)
      bool any = sv->ode.in_flight;
      for (auto& kv : sv->groups) any = any || kv.second->in_flight;
      if (!any) break;
      for (bool done = false; !done;) {
        done = sv->ode.in_flight && event_done(sv->ode.ev);
        for (auto& kv : sv->groups) {
          done = done || (kv.second->in_flight && event_done(kv.second->ev));
        }
        if (!done) std::this_thread::yield();
      }
    }
    size_t active = sv->ode.launched.size();
    for (auto& kv : sv->groups) active += kv.second->members.size();
    if (n_active) *n_active = active;
    if (n_queued) *n_queued = sv->queue.size() + sv->ode_queue.size();
    return 0;
  } catch (const std::exception& e) {
    return fail(e, err, errlen);
  }
}

int dust2_server_poll(dust2_server* sv, uint64_t* ids, double* ll, int* status, size_t max,
                      size_t* n_out, char* err, size_t errlen) {
  try {
    check_handle(sv, TAG_SERVER, "server");
    const size_t n = std::min(max, sv->results.size());
    for (size_t i = 0; i < n; ++i) {
      ids[i] = sv->results[i].id;
      ll[i] = sv->results[i].ll;
      status[i] = sv->results[i].status;
    }
    if (n > 0 && err && errlen) {
      // the first error message among the returned results, if any
      for (size_t i = 0; i < n; ++i) {
        if (sv->results[i].status) {
          std::snprintf(err, errlen, "%s", sv->results[i].message.c_str());
          break;
        }
      }
    }
    sv->results.erase(sv->results.begin(), sv->results.begin() + n);
    *n_out = n;
    return 0;
  } catch (const std::exception& e) {
    return fail(e, err, errlen);
  }
}

int dust2_server_memory(dust2_server* sv, size_t* used, size_t* capacity, char* err,
                        size_t errlen) {
  try {
    check_handle(sv, TAG_SERVER, "server");
    *used = sv->mem.used();
    *capacity = sv->mem.capacity();
    return 0;
  } catch (const std::exception& e) {
    return fail(e, err, errlen);
  }
}

void dust2_server_free(dust2_server* sv) {
  if (!handle_ok(sv, TAG_SERVER)) return;
  // launches in flight read the server's memory: let them finish first
  for (auto& kv : sv->groups) {
    if (kv.second->in_flight) (void)hipEventSynchronize(kv.second->ev);
  }
  if (sv->ode.in_flight) (void)hipEventSynchronize(sv->ode.ev);
  delete sv;
}

}  // extern "C"
