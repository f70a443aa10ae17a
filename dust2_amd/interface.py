"""Python host mirror of dust2's R interface for the accelerated path.

Function names, argument meaning and error behaviour follow the reference's
R API (R/interface.R, R/interface-filter.R, R/interface-unfilter.R,
R/interface-likelihood.R); the work is done by the C-ABI library through
``_lib``.  Arrays use the reference's memory order: a state array that R
shapes [state, particle, group] is returned here as numpy shape
(group, particle, state) over the same memory.
"""
import copy
import ctypes as C

import numpy as np

from . import _lib
from ._lib import DustError, call, dptr, f64, szptr, u32ptr, u64ptr

FP64 = 0
FP32 = 1


class Generator:
    """A model class compiled into the library (the analogue of an odin2 /
    dust2 generator, inst/template/dust.R:1-12)."""

    def __init__(self, name, time_type, parameters, defaults, constant, state, n_data=1,
                 sizes=None, precisions=("fp64", "fp32"), interpolate_beta=None,
                 normalise=None, events=None, state_of=None):
        # state_of: the state names of a parameter row, for models whose
        # state length is a parameter (walk's len); None: `state` always
        # interpolate_beta: None, or the interpolation kind of a beta given
        # as knots 'beta_t' and values 'beta_y' (array parameters)
        self.interpolate_beta = interpolate_beta
        # normalise: maps one group's parameter dict to the parameter
        # layout (array lengths fixed); events: names of the events of a
        # row (a function of the row)
        self.normalise = normalise
        self.events = events
        self.name = name
        self.time_type = time_type
        self.parameters = parameters
        self.defaults = defaults
        self.constant = constant
        self.state = state
        self.n_data = n_data
        self.sizes = sizes or {}
        self.precisions = precisions
        self.state_of = state_of

    def n_state_of(self, row):
        """States of one parameter row."""
        return self.n_state if self.state_of is None else len(self.state_of(row))

    def size(self, name):
        return self.sizes.get(name, 1)

    @property
    def n_pars(self):
        return sum(self.size(k) for k in self.parameters)

    @property
    def n_state(self):
        return len(self.state)

    def __repr__(self):
        return f"<dust_system_generator: {self.name} ({self.time_type})>"


def sir():
    """Stochastic SIR (inst/examples/sir.cpp:12-199)."""
    return Generator(
        "sir", "discrete", ["N", "I0", "beta", "gamma", "exp_noise"],
        {"N": 1000.0, "I0": 10.0, "beta": 0.2, "gamma": 0.1, "exp_noise": 1e6},
        {"N", "exp_noise"}, ["S", "I", "R", "cases_cumul", "cases_inc"])


def sir_time_varying(kind="linear"):
    """The stochastic SIR with beta interpolated in time ('constant',
    'linear' or 'spline'): the model odin2 generates for
    `beta <- interpolate(beta_t, beta_y, kind)` in inst/examples/sir.cpp,
    with array parameters 'beta_t' (knot times, shared by all groups) and
    'beta_y' (values, per group); each step uses beta at its start time
    (dust2/interpolate/interpolate.hpp).  Filters only, pipeline engine."""
    if kind not in INTERP_KINDS:
        raise ValueError(f"Unknown interpolation kind '{kind}'")
    return Generator(
        "sir", "discrete", ["N", "I0", "beta", "gamma", "exp_noise"],
        {"N": 1000.0, "I0": 10.0, "beta": 0.0, "gamma": 0.1, "exp_noise": 1e6},
        {"N", "exp_noise"}, ["S", "I", "R", "cases_cumul", "cases_inc"],
        interpolate_beta=kind)


def _split_beta(generator, pars):
    """Scalar parameters and the (beta_t, beta_y per entry) arrays of a
    generator with an interpolated beta."""
    if generator.interpolate_beta is None or pars is None:
        return pars, None
    single = isinstance(pars, dict)
    plist = [pars] if single else list(pars)
    scal, arrays = [], []
    for p in plist:
        q = dict(p)
        bt, by = q.pop("beta_t", None), q.pop("beta_y", None)
        if (bt is None) != (by is None):
            raise ValueError("'beta_t' and 'beta_y' must be given together")
        scal.append(q)
        arrays.append(None if bt is None else (f64(bt).reshape(-1), f64(by).reshape(-1)))
    return (scal[0] if single else scal), arrays


def _set_beta(obj, arrays, groups):
    """Store the interpolated beta of `groups` (None: all) and upload it."""
    if arrays is None or all(a is None for a in arrays):
        return
    G = obj.n_groups
    groups = list(range(G)) if groups is None else [int(g) for g in groups]
    if len(arrays) == 1 and len(groups) > 1:
        arrays = arrays * len(groups)
    for g, a in zip(groups, arrays):
        if a is None:
            continue
        bt, by = a
        if obj.beta_t is None:
            obj.beta_t = bt.copy()
            obj.beta_y = np.full((G, bt.size), np.nan)
        if bt.size != obj.beta_t.size or not np.array_equal(bt, obj.beta_t):
            raise ValueError("'beta_t' must be the same for every group")
        if by.size != bt.size:
            raise ValueError("Time variable 'beta_t' and interpolation target 'beta_y' must "
                             f"have the same length, but do not ({bt.size} vs {by.size})")
        obj.beta_y[g] = by
    if np.isnan(obj.beta_y).any():
        raise ValueError("'beta_t' and 'beta_y' must be given for every group")
    shared = bool(np.all(obj.beta_y == obj.beta_y[:1]))
    y = np.ascontiguousarray(obj.beta_y[:1] if shared else obj.beta_y)
    call("dust2_filter_set_beta_schedule", obj.ptr, INTERP_KINDS[obj.generator.interpolate_beta],
         dptr(obj.beta_t), obj.beta_t.size, dptr(y), y.shape[0])


def walk():
    """Random walk (inst/examples/walk.cpp): 'len' walkers per particle
    (1..8, constant), each x <- Normal(x, sd * dt) per step; the initial
    state is Normal(0, sd) per walker when 'random_initial', else 0.  A
    system only: the model has no compare_data."""
    return Generator(
        "walk", "discrete", ["sd", "len", "random_initial"],
        {"len": 1.0, "random_initial": 0.0},
        {"len", "random_initial"}, ["x"], n_data=0,
        state_of=lambda row: ["x"] * int(row[1]))


def sirode():
    """Deterministic SIR ODE (inst/examples/sirode.cpp:9-102)."""
    return Generator(
        "sirode", "continuous", ["N", "I0", "beta", "gamma"],
        {"N": 1000.0, "I0": 10.0, "beta": 0.2, "gamma": 0.1},
        {"N"}, ["S", "I", "R", "cases_cumul", "cases_inc"])


def seir16():
    """Age-structured stochastic SEIR, 16 age groups (no counterpart in the
    reference; written to the same model contract).  N: population of each
    age group (16); C: 16 x 16 contact matrix; I0 infectious seeded in every
    age group; weekly incidence reset, Poisson observation of each age
    group's incidence.  FP32 (the force of infection on the matrix cores)
    by default, FP64 in the CPU's arithmetic order when asked."""
    ages = range(16)
    state = ([f"S{a}" for a in ages] + [f"E{a}" for a in ages] + [f"I{a}" for a in ages]
             + [f"R{a}" for a in ages] + [f"inc{a}" for a in ages])
    return Generator(
        "seir16", "discrete", ["beta", "sigma", "gamma", "exp_noise", "I0", "N", "C"],
        {"beta": 0.2, "sigma": 0.25, "gamma": 0.1, "exp_noise": 1e6, "I0": 10.0,
         "N": np.full(16, 1e5), "C": np.full((16, 16), 1.0 / 16)},
        {"N", "C", "exp_noise"}, state, n_data=16, sizes={"N": 16, "C": 256},
        precisions=("fp32", "fp64"))


# ---- continuous-time models with events, carried variables and delays ----
# (dust2_amd/csrc/ode_models.hpp; the reference's example systems)
MAX_CHANGES = 16


def bounce():
    """A ball bouncing with damping: one root event (tests/testthat/examples/
    event.cpp).  'g' is fixed at 9.8 there; here it is a parameter with that
    default."""
    return Generator("bounce", "continuous", ["g", "height", "velocity", "damp"],
                     {"g": 9.8, "height": 0.0, "velocity": 10.0, "damp": 0.9},
                     {"g", "height", "velocity"}, ["height", "velocity"],
                     precisions=("fp64",), events=lambda row: ["bounce"])


def _change_normalise(p):
    # build_shared (tests/testthat/examples/event-time.cpp): r, n, then
    # t_change and delta of length n (padded to the fixed layout)
    out = dict(p)
    if "r" not in p:
        raise ValueError("Expected a value for 'r'")
    n = int(p.get("n", np.size(p.get("t_change", []))))
    tc = f64(p.get("t_change", np.zeros(0))).reshape(-1)
    dl = f64(p.get("delta", np.zeros(0))).reshape(-1)
    if tc.size != n or dl.size != n:
        raise ValueError(f"Expected 't_change' and 'delta' to have length {n}")
    if n > MAX_CHANGES:
        raise ValueError(f"At most {MAX_CHANGES} changes are supported")
    out["n"] = float(n)
    out["t_change"] = np.concatenate([tc, np.full(MAX_CHANGES - n, np.nan)])
    out["delta"] = np.concatenate([dl, np.zeros(MAX_CHANGES - n)])
    return out


def change():
    """Linear growth with jumps at scheduled times: one event per change
    time (tests/testthat/examples/event-time.cpp)."""
    return Generator("change", "continuous", ["r", "n", "t_change", "delta"],
                     {"r": np.nan, "n": 0.0, "t_change": np.full(MAX_CHANGES, np.nan),
                      "delta": np.zeros(MAX_CHANGES)},
                     {"r", "n", "t_change", "delta"}, ["y"],
                     sizes={"t_change": MAX_CHANGES, "delta": MAX_CHANGES},
                     precisions=("fp64",), normalise=_change_normalise,
                     events=lambda row: [f"event-{i + 1}" for i in range(int(row[1]))])


def eventzero():
    """Two accumulating variables, the second reset every period_z, with
    events that only record a crossing (tests/testthat/examples/
    event-zero.cpp)."""
    return Generator("eventzero", "continuous", ["r", "value_y", "value_z", "period_z"],
                     {"r": 1.0, "value_y": 9999.0, "value_z": 9999.0, "period_z": 1.0},
                     {"r", "value_y", "value_z", "period_z"}, ["y", "z"],
                     precisions=("fp64",), events=lambda row: ["y", "z"])


def specialevent():
    """Growth at r0 until an event at time_activate sets the carried
    variable 'active', after which the rate is r1 (tests/testthat/examples/
    specialevent.cpp)."""
    return Generator("specialevent", "continuous", ["r0", "r1", "time_activate"],
                     {"r0": 1.0, "r1": -1.0, "time_activate": np.nan},
                     {"r0", "r1", "time_activate"}, ["y", "active", "end"],
                     precisions=("fp64",),
                     events=lambda row: [] if np.isnan(row[2]) else ["time_activate"])


def sirdelay():
    """The SIR ODE with incidence as the cumulative count less its value one
    time unit earlier, read through a delay (tests/testthat/examples/
    delay.cpp)."""
    return Generator("sirdelay", "continuous", ["N", "I0", "beta", "gamma"],
                     {"N": 1000.0, "I0": 10.0, "beta": 0.2, "gamma": 0.1},
                     {"N"}, ["S", "I", "R", "cases_cumul", "cases_inc"],
                     precisions=("fp64",), events=lambda row: [])


def logdelay():
    """The delayed logistic equation dy/dt = r y(t) (1 - y(t - tau) / K),
    whose derivative reads a delayed value (no reference model; exercises
    delays in the rhs, dust2/continuous/system.hpp:371-387)."""
    return Generator("logdelay", "continuous", ["r", "K", "tau", "y0"],
                     {"r": 1.0, "K": 100.0, "tau": 1.0, "y0": 10.0},
                     {"tau"}, ["y"], precisions=("fp64",), events=lambda row: [])


def special():
    """Mixed time: a is the integral of a carried rate b, and with dt every
    update moves b by a normal step of sd (tests/testthat/examples/
    special.cpp)."""
    return Generator("special", "mixed", ["sd"], {"sd": 1.0}, {"sd"}, ["a", "b"],
                     precisions=("fp64",), events=lambda row: [])


MALARIA_RATES = 15


def _malaria_normalise(p):
    # build_shared (inst/examples/malaria.cpp:160-176): bh, bv and the
    # vector death rate mu = -log(0.9) are fixed
    out = dict(p)
    if int(out.get("n_rates", MALARIA_RATES)) != MALARIA_RATES:
        raise ValueError(f"'n_rates' must be {MALARIA_RATES} (fixed at build time)")
    out["n_rates"] = float(MALARIA_RATES)
    out.setdefault("bh", 0.05)
    out.setdefault("bv", 0.05)
    out.setdefault("mu", -np.log(0.9))
    return out


def malaria():
    """Mixed time: host-vector malaria whose vector birth rate beta follows a
    log-normal random walk with every dt, compared to binomial prevalence
    data (inst/examples/malaria.cpp)."""
    state = (["Sh", "Ih", "Sv", "Iv"] + [f"Ev[{i + 1}]" for i in range(MALARIA_RATES)]
             + ["beta"])
    return Generator("malaria", "mixed",
                     ["a", "bh", "bv", "n_rates", "mu", "r", "tau", "initial_Ih",
                      "initial_Iv", "initial_Sv", "beta_volatility"],
                     {"a": 1.0 / 3.0, "bh": 0.05, "bv": 0.05, "n_rates": 15.0,
                      "mu": -np.log(0.9), "r": 0.01, "tau": 12.0, "initial_Ih": 0.8,
                      "initial_Iv": 18.0, "initial_Sv": 100.0, "beta_volatility": 0.5},
                     {"bh", "bv", "n_rates", "mu", "initial_Ih", "initial_Iv", "initial_Sv",
                      "beta_volatility"}, state, n_data=2, precisions=("fp64",),
                     normalise=_malaria_normalise, events=lambda row: [])


DISCRETE_MODELS = ("sir", "seir16", "walk")
GENERAL_MODELS = ("bounce", "change", "eventzero", "specialevent", "sirdelay", "logdelay",
                  "special", "malaria")


def _precision(p):
    if p in ("fp64", "double", FP64):
        return FP64
    if p in ("fp32", "float", "single", FP32):
        return FP32
    raise ValueError("'precision' must be 'fp64' or 'fp32'")


def pars_matrix(generator, pars, n_groups):
    """Rows of model parameters, one per group; `pars` is a dict (shared by
    all groups) or a list of n_groups dicts (R/interface.R:68-120)."""
    if isinstance(pars, dict):
        pars = [pars] * n_groups
    if len(pars) != n_groups:
        raise ValueError(f"Expected 'pars' to have length {n_groups}, but it had length "
                         f"{len(pars)}")
    rows = []
    for p in pars:
        if generator.normalise is not None:
            p = generator.normalise(p)
        unknown = set(p) - set(generator.parameters)
        if unknown:
            raise ValueError(f"Unknown parameters: {', '.join(sorted(unknown))}")
        row = []
        for k in generator.parameters:
            if k not in p and k not in generator.defaults:
                # a required parameter (walk.cpp's sd, read_real without a default)
                raise ValueError(f"A value is expected for '{k}'")
            v = np.asarray(p[k] if k in p else generator.defaults[k],
                           dtype=np.float64).reshape(-1)
            if v.size != generator.size(k):
                raise ValueError(f"Expected '{k}' to have {generator.size(k)} values")
            row.extend(v.tolist())
        rows.append(row)
    return np.ascontiguousarray(rows, dtype=np.float64)


def _par_offset(generator, name):
    off = 0
    for k in generator.parameters:
        if k == name:
            return off
        off += generator.size(k)
    raise ValueError(f"Unknown parameters: {name}")


def pars_update(generator, pars, current):
    """New parameter rows for an update: names absent from `pars` keep their
    current values (update_shared, inst/examples/sir.cpp:89-93); constant
    parameters cannot change."""
    if isinstance(pars, dict):
        pars = [pars] * current.shape[0]
    if len(pars) != current.shape[0]:
        raise ValueError(f"Expected 'pars' to have length {current.shape[0]}, but it had "
                         f"length {len(pars)}")
    rows = current.copy()
    keys = tuple(pars[0]) if len(pars) else ()
    if all(tuple(p) == keys for p in pars):
        # every row names the same parameters: one array per name
        for k in keys:
            j = _par_offset(generator, k)
            n = generator.size(k)
            v = np.array([p[k] for p in pars], dtype=np.float64).reshape(len(pars), n)
            if k in generator.constant and not np.array_equal(v, rows[:, j:j + n]):
                raise ValueError(f"Parameter '{k}' is constant and cannot be updated")
            rows[:, j:j + n] = v
        return np.ascontiguousarray(rows)
    for i, p in enumerate(pars):
        for k, v in p.items():
            j = _par_offset(generator, k)
            n = generator.size(k)
            v = np.asarray(v, dtype=np.float64).reshape(-1)
            if k in generator.constant and not np.array_equal(v, rows[i, j:j + n]):
                raise ValueError(f"Parameter '{k}' is constant and cannot be updated")
            rows[i, j:j + n] = v
    return np.ascontiguousarray(rows)


def _seed_args(seed):
    """seed: None (random), integer, or raw uint64 stream state (n x 4 words)."""
    if seed is None:
        seed = int(np.random.default_rng().integers(0, 2**31 - 1))
    if isinstance(seed, (int, np.integer)):
        return None, 0, C.c_uint64(int(seed) & (2**64 - 1))
    raw = np.ascontiguousarray(seed, dtype=np.uint64).reshape(-1)
    if raw.size % 4 != 0 or raw.size == 0:
        raise ValueError("Expected raw seed to be a multiple of 4 uint64 words")
    return raw, raw.size // 4, C.c_uint64(0)


def check_index(index, max_value, name, unique=True):
    """R/filter-support.R:73-91 with 0-based indices: every element in
    [0, max_value) and, when asked, distinct."""
    if index is None:
        return None
    idx = np.asarray(index).reshape(-1)
    if idx.size and (not np.all(np.equal(np.mod(idx, 1), 0))):
        raise ValueError(f"'{name}' must be integer-like")
    idx = idx.astype(np.int64)
    if np.any(idx < 0):
        raise ValueError(f"All elements of '{name}' must be at least 0")
    if np.any(idx >= max_value):
        raise ValueError(f"All elements of '{name}' must be at most {max_value - 1}")
    if unique and np.unique(idx).size != idx.size:
        raise ValueError(f"All elements of '{name}' must be distinct")
    return idx


def _index_group(index_group, n_groups):
    if index_group is None:
        return None, 0, n_groups
    idx = np.ascontiguousarray(index_group, dtype=np.uintp).reshape(-1)
    return idx, idx.size, idx.size


class _SaveArgs:
    """save_snapshots and index_state as the run entry points take them
    (R/interface-likelihood.R:83-98 checks them there): data times at which
    the whole state is kept and the (0-based) states the trajectories keep;
    `args` are the four pointer/length arguments of the C call."""

    def __init__(self, save_snapshots, index_state):
        self.snapshots = None
        if save_snapshots is not None and np.size(save_snapshots) > 0:
            self.snapshots = f64(save_snapshots).reshape(-1)
        self.states = None
        if index_state is not None:
            self.states = np.ascontiguousarray(index_state, dtype=np.uintp).reshape(-1)
        n_snap = 0 if self.snapshots is None else self.snapshots.size
        n_states = 0 if self.states is None else self.states.size
        self.args = (dptr(self.snapshots), n_snap, szptr(self.states), n_states)


def _traj_width(obj):
    states = getattr(obj, "last_index_state", None)
    return obj.generator.n_state if states is None else np.size(states)


def _r_num(x):
    """A number as R's cli prints it in messages (6, 0.25)."""
    return f"{float(x):g}"


def check_time_start(time_start, time):
    """R/filter-support.R:94-104."""
    if float(time_start) > float(time[0]):
        raise ValueError(f"'time_start' ({_r_num(time_start)}) is later than the first time "
                         f"in 'data' ({_r_num(time[0])})")


def check_data_groups(n_groups_data, n_groups, shared_data):
    """R/data.R:127-140: shared_data = True takes one group of data for
    every parameter group, False one per group; None (this mirror's
    default) accepts either."""
    if shared_data is None:
        return
    expected = 1 if shared_data else n_groups
    if n_groups_data != expected:
        raise ValueError(f"Expected 'data' to have {expected} groups, but it had "
                         f"{n_groups_data}")


def prepare_data(data, n_groups, name="incidence", n_data=1, shared_data=None):
    """Data as in R/data.R:114-175: a dict with 'time' (n_time) and the
    observed values, shape (n_time,) / (n_time, n_data) shared by all groups,
    or (n_time, n_groups) / (n_time, n_groups, n_data)."""
    out = _prepare_data(data, n_groups, name, n_data)
    check_data_groups(out[2], n_groups, shared_data)
    return out


def _prepare_data(data, n_groups, name, n_data):
    time = f64(data["time"]).reshape(-1)
    nt = time.shape[0]
    obs = f64(data[name])
    if n_data == 1:
        if obs.ndim == 2 and obs.shape[1] == 1:
            obs = obs[:, 0]
        if obs.shape == (nt,):
            return time, np.ascontiguousarray(obs.reshape(nt, 1)), 1
        if obs.shape == (nt, n_groups):
            return time, np.ascontiguousarray(obs), n_groups
    else:
        if obs.shape == (nt, n_data):
            return time, np.ascontiguousarray(obs.reshape(nt, 1, n_data)), 1
        if obs.shape == (nt, n_groups, n_data):
            return time, np.ascontiguousarray(obs), n_groups
    per_group = obs.ndim == (2 if n_data == 1 else 3) and obs.shape[0] == nt and \
        (n_data == 1 or obs.shape[2] == n_data)
    if per_group:
        # R/data.R:134-139
        raise ValueError(f"Expected 'data' to have {n_groups} groups, but it had {obs.shape[1]}")
    raise ValueError(f"Expected '{name}' to have {nt} rows of 1 or {n_groups} groups"
                     + (f" x {n_data} values" if n_data > 1 else ""))


# ---------------------------------------------------------------------------
# Filter
# ---------------------------------------------------------------------------
class Filter:
    """A particle filter object (R/interface-filter.R:37-97), initialised
    lazily on the first likelihood run (R/interface-likelihood.R:48-119)."""

    def __init__(self, generator, time_start, data, n_particles, n_groups, dt, seed,
                 precision, shared_data=None):
        if generator.time_type != "discrete":
            raise ValueError("Can't use 'dust_filter_create()' with continuous-time models")
        self.generator = generator
        self.time_start = float(time_start)
        self.time, self.data, self.n_groups_data = prepare_data(
            data, n_groups, n_data=generator.n_data, shared_data=shared_data)
        check_time_start(time_start, self.time)
        self.n_particles = int(n_particles)
        self.n_groups = int(n_groups)
        self.dt = float(dt)
        # a seed drawn at creation, so that the state reported before the
        # first run is the one the run starts from
        if seed is None:
            seed = int(np.random.default_rng().integers(0, 2**31 - 1))
        self.seed = seed
        if precision is None or (precision == "fp64" and "fp64" not in generator.precisions):
            precision = generator.precisions[0]
        self.precision = _precision(precision)
        self.ptr = None
        self.launch = None
        self.engine = None
        self.chains = None
        self.resample_block = None
        self.beta_t = None
        self.beta_y = None
        self.last_index = None
        self.last_index_state = None
        self.last_snapshots = None

    @property
    def initialised(self):
        return self.ptr is not None

    def _alloc(self, pars):
        pars, arrays = _split_beta(self.generator, pars)
        if self.generator.interpolate_beta is not None and (
                arrays is None or any(a is None for a in arrays)):
            raise ValueError("'beta_t' and 'beta_y' are required by an interpolated beta")
        rows = pars_matrix(self.generator, pars, self.n_groups)
        self.rows = rows
        raw, n_raw, seed = _seed_args(self.seed)
        out = C.c_void_p()
        call("dust2_filter_discrete_alloc", self.generator.name.encode(), dptr(rows),
             self.n_groups, self.n_particles,
             self.time_start, dptr(self.time), self.time.size, self.dt, dptr(self.data),
             self.n_groups_data, u64ptr(raw), n_raw, seed, self.precision, C.byref(out))
        self.ptr = out
        _set_beta(self, arrays, None)
        if self.launch is not None:
            call("dust2_filter_set_launch", self.ptr, *self.launch)
        if self.engine is not None:
            call("dust2_filter_set_engine", self.ptr, *self.engine)
        if self.chains is not None:
            call("dust2_filter_set_pipeline_chains", self.ptr, self.chains)
        if self.resample_block is not None:
            call("dust2_filter_set_pipeline_resample", self.ptr, self.resample_block)

    def set_launch(self, block, particles_per_thread):
        self.launch = (int(block), int(particles_per_thread))
        if self.ptr is not None:
            call("dust2_filter_set_launch", self.ptr, *self.launch)

    ENGINES = {"auto": 0, "persistent": 1, "pipeline": 2}

    def set_engine(self, engine="auto", graph=True, profile=False, chains=None,
                   resample_block=None):
        """Choose the filter engine ('auto', 'persistent', 'pipeline'),
        HIP-graph replay of the pipeline, per-kernel profiling, the
        pipeline's number of launch chains (1-4) and its resample workgroup
        width (1024, 512 or 256 threads)."""
        self.engine = (self.ENGINES[engine], int(bool(graph)), int(bool(profile)))
        if chains is not None:
            self.chains = int(chains)
        if resample_block is not None:
            self.resample_block = int(resample_block)
        if self.ptr is not None:
            call("dust2_filter_set_engine", self.ptr, *self.engine)
            if self.chains is not None:
                call("dust2_filter_set_pipeline_chains", self.ptr, self.chains)
            if self.resample_block is not None:
                call("dust2_filter_set_pipeline_resample", self.ptr, self.resample_block)

    def autotune(self, pars=None, index_group=None):
        """Time every engine / launch shape on this workload and keep the
        fastest; returns the candidates as dicts: persistent (block, ppt,
        ms) and pipeline (chains, resample_block, ms)."""
        if self.ptr is None:
            if pars is None:
                raise ValueError("'pars' cannot be NULL, as 'obj' is not initialised")
            self._alloc(pars)
        elif pars is not None:
            idx, n_idx, _ = _index_group(index_group, self.n_groups)
            sel = np.arange(self.n_groups) if idx is None else idx.astype(np.int64)
            rows = pars_update(self.generator, pars, self.rows[sel])
            call("dust2_filter_update_pars", self.ptr, dptr(rows), szptr(idx), n_idx)
            self.rows[sel] = rows
        idx, n_idx, _ = _index_group(index_group, self.n_groups)
        rep = np.zeros((32, 4))
        n = np.zeros(1, dtype=np.uintp)
        call("dust2_filter_autotune", self.ptr, szptr(idx), n_idx, dptr(rep), 32, szptr(n))
        names = {1: "persistent", 2: "pipeline"}
        out = []
        for r in rep[:int(n[0])]:
            e = names[int(r[0])]
            if e == "persistent":
                out.append({"engine": e, "block": int(r[1]), "ppt": int(r[2]), "ms": float(r[3])})
            else:
                out.append({"engine": e, "chains": int(r[2]), "resample_block": int(r[1]),
                            "ms": float(r[3])})
        best = min(out, key=lambda r: r["ms"])
        self.engine = (self.ENGINES[best["engine"]], 1, 0)
        if best["engine"] == "persistent":
            self.launch = (best["block"], best["ppt"])
        else:
            self.chains = best["chains"]
            self.resample_block = best["resample_block"]
        return out

    def __del__(self):
        if getattr(self, "ptr", None) is not None and _lib._lib is not None:
            _lib.lib().dust2_filter_free(self.ptr)
            self.ptr = None


class ModelFilter:
    """A particle filter over a mixed-time model (the general ODE engine):
    dust_filter_create for a model run in continuous time with discrete
    updates (R/interface-filter.R:37-97, tests/testthat/test-filter.R:705-722),
    over dust2/filter.hpp.  Data fields per model (malaria: tested,
    positive); FP64."""

    DATA_FIELDS = {"malaria": ["tested", "positive"]}

    def __init__(self, generator, time_start, data, n_particles, n_groups, dt, seed,
                 ode_control=None, shared_data=None):
        self.generator = generator
        self.time_start = float(time_start)
        self.n_particles = int(n_particles)
        self.n_groups = int(n_groups)
        self.dt = 0.0 if dt is None else float(dt)
        fields = self.DATA_FIELDS.get(generator.name)
        if fields is None:
            raise ValueError(f"'{generator.name}' has no compare_data")
        self.time = np.ascontiguousarray(f64(data["time"]).reshape(-1))
        cols = [np.asarray(f64(data[k])) for k in fields]
        cols = [c.reshape(self.time.size, -1) for c in cols]
        self.n_groups_data = cols[0].shape[1]
        check_data_groups(self.n_groups_data, self.n_groups, shared_data)
        if self.n_groups_data not in (1, self.n_groups):
            raise ValueError(f"Expected 'data' to have {self.n_groups} groups, but it had "
                             f"{self.n_groups_data}")
        check_time_start(time_start, self.time)
        # [time][group][field]
        self.data = np.ascontiguousarray(np.stack(cols, axis=2))
        if seed is None:
            seed = int(np.random.default_rng().integers(0, 2**31 - 1))
        self.seed = seed
        self.ode_control = _ode_control(ode_control)
        self.ptr = None
        self.rows = None
        self.last_index = None
        self.last_index_state = None
        self.last_snapshots = None

    @property
    def initialised(self):
        return self.ptr is not None

    def _alloc(self, pars):
        self.rows = pars_matrix(self.generator, pars, self.n_groups)
        raw, n_raw, seed = _seed_args(self.seed)
        out = C.c_void_p()
        call("dust2_filter_ode_alloc", self.generator.name.encode(), dptr(self.rows),
             self.n_groups, self.n_particles, self.time_start, dptr(self.time), self.time.size,
             dptr(self.data), self.n_groups_data, self.dt, dptr(self.ode_control.values),
             u64ptr(raw), n_raw, seed, C.byref(out))
        self.ptr = out

    def _update(self, pars, index_group):
        """Parameters for the groups of index_group (R/interface-likelihood.R:
        54-75); returns the index array for the C call."""
        idx, n_idx, n_out = _index_group(index_group, self.n_groups)
        if self.ptr is None:
            if pars is None:
                raise ValueError("'pars' cannot be NULL, as 'obj' is not initialised")
            if index_group is not None:
                raise ValueError("'index_group' must be NULL, as 'obj' is not initialised")
            self._alloc(pars)
        elif pars is not None:
            sel = np.arange(self.n_groups) if idx is None else idx.astype(np.int64)
            rows = self.rows.copy()
            rows[sel] = pars_update(self.generator, pars, self.rows[sel])
            call(self.prefix + "_update_pars", self.ptr, dptr(rows))
            self.rows = rows
        return idx, n_idx, n_out

    def _initial(self, initial):
        if initial is None:
            return None
        return np.broadcast_to(
            f64(initial), (self.n_groups, self.n_particles, self.generator.n_state)).copy()

    def run(self, pars, index_group=None, initial=None, save_trajectories=False,
            save_snapshots=None, index_state=None):
        idx, n_idx, n_out = self._update(pars, index_group)
        save = _SaveArgs(save_snapshots, index_state)
        ll = np.zeros(n_out)
        self.last_snapshots = None
        call("dust2_filter_ode_run", self.ptr, dptr(self._initial(initial)),
             int(bool(save_trajectories)), *save.args, szptr(idx), n_idx, dptr(ll))
        self._saved(index_group, save)
        return float(ll[0]) if (self.n_groups == 1 and index_group is None) else ll

    def _saved(self, index_group, save):
        self.last_index = index_group
        self.last_index_state = save.states
        self.last_snapshots = save.snapshots

    def last_state(self):
        if self.ptr is None:
            raise DustError("System has not been run")
        out = np.zeros((self.n_groups, self.n_particles, self.generator.n_state))
        call(self.prefix + "_last_state", self.ptr, dptr(out))
        return out

    def last_state_selected(self):
        """The chosen particle's state for each group of the last run
        (filter<T>::select_random_particle, dust2/filter.hpp:130-138)."""
        if self.ptr is None:
            raise DustError("State is not current")
        n = self.n_groups if self.last_index is None else np.size(self.last_index)
        out = np.zeros((n, self.generator.n_state))
        call("dust2_filter_ode_last_state_selected", self.ptr, dptr(out))
        return out

    def _select_args(self, select_random_particle):
        return (int(bool(select_random_particle)),)

    def last_trajectories(self, select_random_particle=False):
        if self.ptr is None:
            raise DustError("Trajectories are not current")
        n = self.n_groups if self.last_index is None else np.size(self.last_index)
        p = 1 if select_random_particle else self.n_particles
        out = np.zeros((self.time.size, n, p, _traj_width(self)))
        call(self.prefix + "_last_trajectories", self.ptr,
             *self._select_args(select_random_particle), dptr(out))
        return out

    def last_snapshots_state(self, select_random_particle=False):
        if self.ptr is None or self.last_snapshots is None:
            raise DustError("Snapshots are not current")
        n = self.n_groups if self.last_index is None else np.size(self.last_index)
        p = 1 if select_random_particle else self.n_particles
        out = np.zeros((self.last_snapshots.size, n, p, self.generator.n_state))
        call(self.prefix + "_last_snapshots", self.ptr,
             *self._select_args(select_random_particle), dptr(out))
        return out

    prefix = "dust2_filter_ode"

    def __del__(self):
        if getattr(self, "ptr", None) is not None and _lib._lib is not None:
            getattr(_lib.lib(), self.prefix + "_free")(self.ptr)
            self.ptr = None


class ModelUnfilter(ModelFilter):
    """dust_unfilter_create for a model of the general engine with data
    (R/interface-unfilter.R:20-71 over dust2/unfilter.hpp): deterministic
    updates, one log-likelihood per particle."""
    prefix = "dust2_unfilter_ode"

    def _select_args(self, select_random_particle):
        # every particle of an unfilter is the same deterministic run: the
        # callers take the first
        if select_random_particle:
            raise ValueError("'select_random_particle' is not supported by this filter")
        return ()

    def __init__(self, generator, time_start, data, n_particles, n_groups, dt, ode_control,
                 shared_data=None):
        super().__init__(generator, time_start, data, n_particles, n_groups, dt, 0, ode_control,
                         shared_data)

    def _alloc(self, pars):
        self.rows = pars_matrix(self.generator, pars, self.n_groups)
        out = C.c_void_p()
        call("dust2_unfilter_ode_alloc", self.generator.name.encode(), dptr(self.rows),
             self.n_groups, self.n_particles, self.time_start, dptr(self.time), self.time.size,
             dptr(self.data), self.n_groups_data, self.dt, dptr(self.ode_control.values),
             C.byref(out))
        self.ptr = out

    def run(self, pars, index_group=None, initial=None, save_trajectories=False,
            save_snapshots=None, index_state=None):
        idx, n_idx, n_out = self._update(pars, index_group)
        save = _SaveArgs(save_snapshots, index_state)
        ll = np.zeros((n_out, self.n_particles))
        self.last_snapshots = None
        call("dust2_unfilter_ode_run", self.ptr, dptr(self._initial(initial)),
             int(bool(save_trajectories)), *save.args, szptr(idx), n_idx, dptr(ll))
        self._saved(index_group, save)
        if self.n_particles == 1:
            ll = ll[:, 0]
        return float(ll[0]) if (self.n_groups == 1 and ll.ndim == 1 and index_group is None) \
            else ll


def dust_filter_create(generator, time_start, data, n_particles, n_groups=1, dt=1,
                       seed=None, precision=None, ode_control=None, shared_data=None):
    """R/interface-filter.R:37-97.  shared_data: True for one group of data
    used by every parameter group, False for one group of data per
    parameter group; None (the default here) accepts either."""
    if generator.time_type == "mixed":
        return ModelFilter(generator, time_start, data, n_particles, n_groups, dt, seed,
                           ode_control, shared_data)
    return Filter(generator, time_start, data, n_particles, n_groups, dt, seed, precision,
                  shared_data)


def dust_likelihood_run(obj, pars, initial=None, save_trajectories=False, index_group=None,
                        adjoint=False, save_snapshots=None, index_state=None):
    """Run the filter (or unfilter) and return the log-likelihood per group
    (a float for a single group), R/interface-likelihood.R:48-119.
    index_state (0-based) selects the states kept in the trajectories;
    snapshots and the final state keep every state."""
    index_group = check_index(index_group, obj.n_groups, "index_group")
    if obj.ptr is None:
        if pars is None:
            raise ValueError("'pars' cannot be NULL, as 'obj' is not initialised")
        if index_group is not None:
            raise ValueError("'index_group' must be NULL, as 'obj' is not initialised")
    index_state = check_index(index_state, obj.generator.n_state, "index_state")
    if isinstance(obj, ModelFilter):
        return obj.run(pars, index_group, initial=initial, save_trajectories=save_trajectories,
                       save_snapshots=save_snapshots, index_state=index_state)
    if isinstance(obj, Unfilter):
        return obj.run(pars, index_group, initial=initial, adjoint=adjoint,
                       save_trajectories=save_trajectories, save_snapshots=save_snapshots,
                       index_state=index_state)
    if not obj.initialised:
        obj._alloc(pars)
    elif pars is not None:
        idx, n_idx, _ = _index_group(index_group, obj.n_groups)
        sel = np.arange(obj.n_groups) if idx is None else idx.astype(np.int64)
        pars, arrays = _split_beta(obj.generator, pars)
        rows = pars_update(obj.generator, pars, obj.rows[sel])
        call("dust2_filter_update_pars", obj.ptr, dptr(rows), szptr(idx), n_idx)
        obj.rows[sel] = rows
        _set_beta(obj, arrays, sel)
    idx, n_idx, n_out = _index_group(index_group, obj.n_groups)
    init = None
    if initial is not None:
        init = f64(initial).reshape(-1)
        if init.size != obj.n_groups * obj.n_particles * obj.generator.n_state:
            raise ValueError("Expected 'initial' to have n_groups * n_particles * n_state values")
    ll = np.zeros(n_out)
    save = _SaveArgs(save_snapshots, index_state)
    obj.last_snapshots = None
    call("dust2_filter_run", obj.ptr, dptr(init), int(bool(save_trajectories)), *save.args,
         szptr(idx), n_idx, dptr(ll))
    obj.last_index = index_group
    obj.last_index_state = save.states
    obj.last_snapshots = save.snapshots
    return float(ll[0]) if (obj.n_groups == 1 and index_group is None) else ll


def dust_likelihood_ensure_initialised(obj, pars):
    """R/interface-likelihood.R:121-139: allocate the object's device state
    with `pars` unless it already exists; True when it was allocated now."""
    if obj.ptr is not None:
        return False
    if pars is None:
        raise ValueError("'pars' cannot be NULL, as 'obj' is not initialised")
    obj._alloc(pars)
    return True


def dust_likelihood_copy(obj, seed=None):
    """R/interface-likelihood.R:156-169: an independent copy of a filter or
    unfilter with the same settings.  The copy is uninitialised (its device
    state is allocated on its first run) and a filter copy draws from the
    streams of `seed` (a fresh seed when None), not from the parent's."""
    dst = copy.copy(obj)
    dst.ptr = None
    dst.last_index = None
    dst.last_index_state = None
    dst.last_snapshots = None
    if isinstance(obj, Filter):
        dst.seed = int(np.random.default_rng().integers(0, 2**31 - 1)) if seed is None else seed
    return dst


def _filter_only(obj, what, engine=False):
    """Particle filters only; with engine, only the filters of the discrete
    SIR/SEIR engines (their statistics, ancestry and raw trajectories)."""
    if not isinstance(obj, (Filter, ModelFilter)) or isinstance(obj, ModelUnfilter):
        raise ValueError(f"'{what}' is only available for particle filters")
    if engine and isinstance(obj, ModelFilter):
        raise ValueError(f"'{what}' is not available for filters over continuous-time models")


def dust_likelihood_last_state(obj, select_random_particle=False):
    """Final state (n_groups, n_particles, n_state); with
    select_random_particle, the chosen particle of each group of the last
    run (n_groups, n_state).  For an unfilter every particle is the same
    deterministic run, so the selected particle is the first."""
    if isinstance(obj, Unfilter):
        if obj.ptr is None:
            raise DustError("System has not been run")
        st = obj.last_state()
        return st[:, 0, :] if select_random_particle else st
    if isinstance(obj, ModelFilter):
        if select_random_particle:
            if isinstance(obj, ModelUnfilter):
                return obj.last_state()[:, 0, :]
            return obj.last_state_selected()
        return obj.last_state()
    if select_random_particle:
        n_groups = obj.n_groups if obj.last_index is None else np.size(obj.last_index)
        out = np.zeros((n_groups, obj.generator.n_state))
        call("dust2_filter_last_state_selected", obj.ptr, dptr(out))
        return out
    out = np.zeros((obj.n_groups, obj.n_particles, obj.generator.n_state))
    call("dust2_filter_last_state", obj.ptr, dptr(out))
    return out


def dust_likelihood_last_snapshots(obj, select_random_particle=False):
    """The whole state at the snapshot times of the last run, following each
    exported particle's ancestry (a filter) or as computed (an unfilter,
    dust2/r/unfilter.hpp:97-128): (n_snapshots, n_groups, n_particles or 1,
    n_state)."""
    if isinstance(obj, ModelUnfilter):
        if select_random_particle:
            raise ValueError("'select_random_particle' is not supported by this filter")
        return obj.last_snapshots_state()
    if isinstance(obj, ModelFilter):
        return obj.last_snapshots_state(select_random_particle)
    if isinstance(obj, Unfilter):
        return obj.last_snapshots_state()
    if obj.ptr is None or getattr(obj, "last_snapshots", None) is None:
        raise DustError("Snapshots are not current")
    n_groups = obj.n_groups if obj.last_index is None else np.size(obj.last_index)
    p_out = 1 if select_random_particle else obj.n_particles
    out = np.zeros((obj.last_snapshots.size, n_groups, p_out, obj.generator.n_state))
    call("dust2_filter_last_snapshots", obj.ptr, int(bool(select_random_particle)), dptr(out))
    return out


def dust_likelihood_last_trajectories(obj, select_random_particle=False):
    """dust_likelihood_last_trajectories (R/interface-likelihood.R): for a
    filter each particle's trajectory follows its ancestry back from the
    last time (dust2/trajectories.hpp:128-189); for an unfilter the state at
    every data time.  (n_time, n_groups of the last run, n_particles or 1,
    n_state)."""
    if isinstance(obj, Unfilter):
        out = obj.last_trajectories()
    elif isinstance(obj, ModelFilter):
        out = obj.last_trajectories(select_random_particle)
    else:
        if obj.ptr is None:
            raise DustError("Trajectories are not current")
        n_groups = obj.n_groups if obj.last_index is None else np.size(obj.last_index)
        p_out = 1 if select_random_particle else obj.n_particles
        out = np.zeros((obj.time.size, n_groups, p_out, _traj_width(obj)))
        call("dust2_filter_last_trajectories", obj.ptr, int(bool(select_random_particle)),
             dptr(out))
    return out


def dust_likelihood_last_trajectories_raw(obj):
    """The saved post-resampling states, not following the ancestry."""
    _filter_only(obj, "last_trajectories_raw", engine=True)
    out = np.zeros((obj.time.size, obj.n_groups, obj.n_particles, obj.generator.n_state))
    call("dust2_filter_last_trajectories_raw", obj.ptr, dptr(out))
    return out


def dust_likelihood_last_ancestry(obj):
    _filter_only(obj, "last_ancestry", engine=True)
    out = np.zeros((obj.time.size, obj.n_groups, obj.n_particles), dtype=np.uint32)
    call("dust2_filter_last_ancestry", obj.ptr, u32ptr(out))
    return out


def dust_likelihood_rng_state(obj):
    """R/interface-likelihood.R:281-288: before the first run, the state the
    run will start from (the seed expanded to n_groups * (n_particles + 2)
    streams, dust2/r/discrete/filter.hpp:63-74)."""
    _filter_only(obj, "rng_state")
    n = obj.n_groups * (obj.n_particles + 2)
    if obj.ptr is None:
        raw, n_raw, seed = _seed_args(obj.seed)
        out = np.zeros(n * 4, dtype=np.uint64)
        call("dust2_amd_rng_streams", u64ptr(raw), n_raw, seed, n, u64ptr(out))
        return out
    out = np.zeros(n * 4, dtype=np.uint64)
    name = "dust2_filter_ode_rng_state" if isinstance(obj, ModelFilter) else \
        "dust2_filter_rng_state"
    call(name, obj.ptr, u64ptr(out))
    return out


def dust_likelihood_set_rng_state(obj, state):
    """R/interface-likelihood.R:297-306: before the first run the state
    replaces the seed."""
    _filter_only(obj, "set_rng_state")
    st = np.ascontiguousarray(state, dtype=np.uint64).reshape(-1)
    if st.size != obj.n_groups * (obj.n_particles + 2) * 4:
        raise ValueError("Incorrect length for 'rng_state'")
    if obj.ptr is None:
        obj.seed = st.reshape(-1, 4).copy()
        return
    name = "dust2_filter_ode_set_rng_state" if isinstance(obj, ModelFilter) else \
        "dust2_filter_set_rng_state"
    call(name, obj.ptr, u64ptr(st))


def dust_likelihood_last_stats(obj):
    _filter_only(obj, "last_stats", engine=True)
    out = np.zeros(7)
    call("dust2_filter_last_stats", obj.ptr, dptr(out))
    engines = {1: "persistent", 2: "pipeline"}
    return {"n_fallback": int(out[0]), "kernel_ms": float(out[1]),
            "engine": engines.get(int(out[2]), "none"), "advance_ms": float(out[3]),
            "resample_ms": float(out[4]), "launch": (int(out[5]), int(out[6]))}


# ---------------------------------------------------------------------------
# System
# ---------------------------------------------------------------------------
class System:
    """A discrete-time system (dust_system_create, R/interface.R:60-150) of
    any discrete-time model compiled into the library: one set of kernels
    templated on the model, the model named at allocation."""
    prefix = "dust2_system_discrete"

    def __init__(self, generator, pars, n_particles, n_groups, time, dt, seed, deterministic,
                 precision):
        if generator.name not in DISCRETE_MODELS or generator.interpolate_beta is not None:
            raise ValueError(f"No system kernels for '{generator.name}'")
        self.generator = generator
        self.n_particles = int(n_particles)
        self.n_groups = int(n_groups)
        rows = pars_matrix(generator, pars, self.n_groups)
        self.rows = rows
        self.n_state = generator.n_state_of(rows[0])
        if precision is None:
            precision = generator.precisions[0]
        raw, n_raw, s = _seed_args(seed)
        out = C.c_void_p()
        call("dust2_system_discrete_alloc", generator.name.encode(), dptr(rows), self.n_groups,
             self.n_particles, float(time), float(dt), u64ptr(raw), n_raw, s,
             int(bool(deterministic)), _precision(precision), C.byref(out))
        self.ptr = out

    def __del__(self):
        if getattr(self, "ptr", None) is not None and _lib._lib is not None:
            _lib.lib().dust2_system_discrete_free(self.ptr)
            self.ptr = None

    @property
    def shape(self):
        return (self.n_groups, self.n_particles, self.n_state)


def _log_cap(on, n_rows):
    """Entries kept per particle by a per-particle log (step times, history):
    up to 100,000, fewer for many particles so that the log stays within
    about 1 GiB of FP64 words, never below 1,000."""
    if not on:
        return 0
    return int(max(1000, min(100000, (1 << 27) // max(n_rows, 1))))


class ContinuousSystem:
    """A continuous-time system (dust_system_create for an ODE model,
    R/interface.R:60-150; dust2/continuous/system.hpp): FP64, solver
    internals kept on the device between calls."""
    prefix = "dust2_system_sirode"

    def __init__(self, generator, pars, n_particles, n_groups, time, ode_control):
        if generator.name != "sirode":
            raise ValueError(f"No continuous system kernels for '{generator.name}'")
        self.generator = generator
        self.n_particles = int(n_particles)
        self.n_groups = int(n_groups)
        self.rows = pars_matrix(generator, pars, self.n_groups)
        self.ode_control = _ode_control(ode_control)
        ctl = self.ode_control
        self.record = _log_cap(ctl.debug_record_step_times, self.n_groups * self.n_particles)
        out = C.c_void_p()
        call("dust2_system_sirode_alloc", dptr(self.rows), self.n_groups, self.n_particles,
             float(time), dptr(ctl.values), dptr(ctl.critical_times), ctl.critical_times.size,
             self.record, C.byref(out))
        self.ptr = out

    def __del__(self):
        if getattr(self, "ptr", None) is not None and _lib._lib is not None:
            _lib.lib().dust2_system_sirode_free(self.ptr)
            self.ptr = None

    @property
    def shape(self):
        return (self.n_groups, self.n_particles, self.generator.n_state)


class GeneralSystem:
    """A continuous-time system of a model with events, carried variables,
    resets or delays (dust_system_create, R/interface.R:60-150, over
    dust2/continuous/system.hpp): FP64; solver internals, the event log and
    the history ring stay on the device between calls."""
    prefix = "dust2_system_ode"

    def __init__(self, generator, pars, n_particles, n_groups, time, ode_control, dt=None,
                 seed=None, deterministic=False, history_cap=None, event_cap=1024):
        self.generator = generator
        self.n_particles = int(n_particles)
        self.n_groups = int(n_groups)
        sizes = np.zeros(9, dtype=np.uintp)
        call("dust2_system_ode_model", generator.name.encode(), szptr(sizes))
        (self.n_var, self.n_special, self.n_output, n_pars, self.max_events,
         self.n_hist, has_delays, mixed, self.n_data) = (int(v) for v in sizes)
        self.has_delays = bool(has_delays)
        self.mixed = bool(mixed)
        self.dt = 0.0 if dt is None else float(dt)
        self.rows = pars_matrix(generator, pars, self.n_groups)
        if self.rows.shape[1] != n_pars:
            raise ValueError(f"Parameter layout of '{generator.name}' does not match the library")
        self.ode_control = _ode_control(ode_control)
        ctl = self.ode_control
        n_total = self.n_groups * self.n_particles
        self.record = _log_cap(ctl.debug_record_step_times, n_total)
        self.save_history = bool(ctl.save_history) or self.has_delays
        if history_cap is None:
            # every step for inspection, else enough for the delays
            history_cap = _log_cap(True, n_total * (3 + 5 * self.n_hist)) \
                if ctl.save_history else 4096
        self.history_cap = int(history_cap) if self.save_history else 0
        self.event_cap = int(event_cap) if self.max_events else 0
        raw, n_raw, s = _seed_args(seed)
        out = C.c_void_p()
        call("dust2_system_ode_alloc", generator.name.encode(), dptr(self.rows), self.n_groups,
             self.n_particles, float(time), self.dt, dptr(ctl.values),
             dptr(ctl.critical_times), ctl.critical_times.size, self.record,
             int(ctl.save_history), self.history_cap, self.event_cap, u64ptr(raw), n_raw, s,
             int(bool(deterministic)), C.byref(out))
        self.ptr = out

    def __del__(self):
        if getattr(self, "ptr", None) is not None and _lib._lib is not None:
            _lib.lib().dust2_system_ode_free(self.ptr)
            self.ptr = None

    @property
    def shape(self):
        return (self.n_groups, self.n_particles, self.generator.n_state)


def dust_system_create(generator, pars=None, n_particles=1, n_groups=1, time=0, dt=None,
                       seed=None, deterministic=False, precision=None, ode_control=None,
                       **kwargs):
    if pars is None:
        pars = {}
    if generator.time_type == "mixed":
        # dt given: the model's update runs every dt; none: continuous only
        # (R/interface.R:600-620)
        return GeneralSystem(generator, pars, n_particles, n_groups, time, ode_control, dt=dt,
                             seed=seed, deterministic=deterministic, **kwargs)
    if generator.time_type == "continuous":
        if dt is not None:
            raise ValueError("Can't use 'dt' with continuous-time systems")
        if generator.name in GENERAL_MODELS:
            return GeneralSystem(generator, pars, n_particles, n_groups, time, ode_control,
                                 **kwargs)
        return ContinuousSystem(generator, pars, n_particles, n_groups, time, ode_control)
    if ode_control is not None:
        raise ValueError("Can't use 'ode_control' with discrete-time systems")
    return System(generator, pars, n_particles, n_groups, time, 1 if dt is None else dt, seed,
                  deterministic, precision)


def _prefix(sys):
    return sys.prefix


def dust_system_run_to_time(sys, time):
    call(_prefix(sys) + "_run_to_time", sys.ptr, float(time))


def dust_system_simulate(sys, times, index_state=None):
    """The state at each time, (n_times, n_groups, n_particles, n_state);
    index_state (0-based) keeps some states (R/interface.R, simulate)."""
    t = f64(times).reshape(-1)
    out = np.zeros((t.size,) + sys.shape)
    call(_prefix(sys) + "_simulate", sys.ptr, dptr(t), t.size, dptr(out))
    if index_state is not None:
        idx = check_index(index_state, sys.shape[2], "index_state")
        out = np.ascontiguousarray(out[..., idx])
    return out


def dust_system_state(sys):
    out = np.zeros(sys.shape)
    call(_prefix(sys) + "_state", sys.ptr, dptr(out))
    return out


def dust_system_set_state(sys, state):
    st = np.broadcast_to(f64(state), sys.shape).copy()
    call(_prefix(sys) + "_set_state", sys.ptr, dptr(st))


def dust_system_set_state_initial(sys):
    call(_prefix(sys) + "_set_state_initial", sys.ptr)


def dust_system_time(sys):
    out = np.zeros(1)
    call(_prefix(sys) + "_time", sys.ptr, dptr(out))
    return float(out[0])


def dust_system_set_time(sys, time):
    call(_prefix(sys) + "_set_time", sys.ptr, float(time))


def dust_system_update_pars(sys, pars):
    rows = pars_update(sys.generator, pars, sys.rows)
    call(_prefix(sys) + "_update_pars", sys.ptr, dptr(rows))
    sys.rows = rows


def dust_system_compare_data(sys, data):
    """data: one value (shared) or one per group; returns (group, particle).
    For a model with several data fields (malaria: tested, positive), a dict
    of them or an array [group][field]."""
    if isinstance(sys, GeneralSystem):
        if isinstance(data, dict):
            names = ["tested", "positive"]
            d = np.column_stack([np.atleast_1d(f64(data.get(k, np.nan))) for k in names])
        else:
            d = np.atleast_2d(f64(data))
        d = np.ascontiguousarray(d)
        out = np.zeros((sys.n_groups, sys.n_particles))
        call("dust2_system_ode_compare_data", sys.ptr, dptr(d), d.shape[0], dptr(out))
        return out
    nd = max(getattr(sys.generator, "n_data", 1), 1)
    d = f64(data).reshape(-1)
    if d.size % nd != 0 or d.size // nd not in (1, sys.n_groups):
        # R/interface.R (dust_system_compare_data)
        raise ValueError(f"Expected 'data' to have length {sys.n_groups}, but it had length "
                         f"{d.size // nd}")
    out = np.zeros((sys.n_groups, sys.n_particles))
    call(_prefix(sys) + "_compare_data", sys.ptr, dptr(d), d.size // nd, dptr(out))
    return out


def dust_system_reorder(sys, index):
    idx = np.ascontiguousarray(np.broadcast_to(index, (sys.n_groups, sys.n_particles)),
                               dtype=np.uint32)
    call(_prefix(sys) + "_reorder", sys.ptr, u32ptr(idx))


def dust_system_internals(sys, include_history=False):
    """dust_system_internals for a continuous-time system (dust2/r/continuous/
    system.hpp:56-160): per particle (group-major) n_steps,
    n_steps_accepted, n_steps_rejected, step_size; with
    debug_record_step_times, the step times; for models with events, the
    events found (name, time, 1-based index, sign); with include_history,
    the saved history (t0, t1, h and coefficients [variable, 5, step])."""
    if not isinstance(sys, (ContinuousSystem, GeneralSystem)):
        raise ValueError("Internals are only available for continuous-time systems")
    n = sys.n_groups * sys.n_particles
    stats = np.zeros((n, 4))
    counts = np.zeros(n, dtype=np.uintp)
    times = np.zeros(n * max(sys.record, 1))
    call(sys.prefix + "_internals", sys.ptr, dptr(stats),
         dptr(times) if sys.record else None, szptr(counts))
    events = history = None
    if isinstance(sys, GeneralSystem):
        if sys.event_cap:
            ev_n = np.zeros(n, dtype=np.uintp)
            ev = np.zeros(n * sys.event_cap * 3)
            call("dust2_system_ode_events", sys.ptr, szptr(ev_n), dptr(ev))
            events = (ev_n, ev.reshape(n, sys.event_cap, 3))
        if include_history and sys.history_cap:
            h_n = np.zeros(n, dtype=np.uintp)
            hw = 3 + 5 * sys.n_hist
            hist = np.zeros(n * sys.history_cap * hw)
            call("dust2_system_ode_history", sys.ptr, szptr(h_n), dptr(hist))
            history = (h_n, hist.reshape(n, sys.history_cap, hw))
    out = []
    for k in range(n):
        row = {"group": k // sys.n_particles, "particle": k % sys.n_particles,
               "n_steps": int(stats[k, 0]), "n_steps_accepted": int(stats[k, 1]),
               "n_steps_rejected": int(stats[k, 2]), "step_size": stats[k, 3]}
        if sys.record:
            c = min(int(counts[k]), sys.record)
            row["step_times"] = times[k * sys.record:k * sys.record + c].copy()
        if events is not None:
            names = sys.generator.events(sys.rows[k // sys.n_particles])
            c = int(events[0][k])
            if c > sys.event_cap:
                raise RuntimeError(f"{c} events found but only {sys.event_cap} kept; "
                                   "raise 'event_cap'")
            r = events[1][k, :c]
            idx = r[:, 1].astype(int)
            row["events"] = {"name": [names[i] for i in idx], "time": r[:, 0].copy(),
                             "index": idx + 1, "sign": r[:, 2].copy()}
        if history is not None:
            c = int(history[0][k])
            kept = min(c, sys.history_cap)
            e = history[1][k, :kept]
            H = sys.n_hist
            row["history"] = {"t0": e[:, 0].copy(), "t1": e[:, 1].copy(), "h": e[:, 2].copy(),
                              "coefficients": e[:, 3:].reshape(kept, 5, H).transpose(2, 1, 0),
                              "n_saved": c}
        out.append(row)
    return out


def dust_system_rng_state(sys):
    out = np.zeros(sys.n_groups * sys.n_particles * 4, dtype=np.uint64)
    name = "dust2_system_ode_rng_state" if isinstance(sys, GeneralSystem) else \
        "dust2_system_discrete_rng_state"
    call(name, sys.ptr, u64ptr(out))
    return out


def dust_system_set_rng_state(sys, state):
    st = np.ascontiguousarray(state, dtype=np.uint64).reshape(-1)
    name = "dust2_system_ode_set_rng_state" if isinstance(sys, GeneralSystem) else \
        "dust2_system_discrete_set_rng_state"
    call(name, sys.ptr, u64ptr(st))


# ---------------------------------------------------------------------------
# Unfilter (deterministic likelihood)
# ---------------------------------------------------------------------------
class OdeControl:
    """dust_ode_control (R/interface-continuous.R:51-79): the five solver
    settings (as an array: max_steps, atol, rtol, step_size_min,
    step_size_max), the times the solver must stop at, whether systems
    record their step times and whether they save the history of their
    steps (models with delays always do)."""

    def __init__(self, values, critical_times=None, debug_record_step_times=False,
                 save_history=False):
        self.values = f64(values).reshape(-1)
        self.save_history = bool(save_history)
        self.critical_times = (np.zeros(0) if critical_times is None
                               else np.ascontiguousarray(f64(critical_times).reshape(-1)))
        if np.any(np.diff(self.critical_times) <= 0):
            raise ValueError("'critical_times' must be strictly increasing")
        self.debug_record_step_times = bool(debug_record_step_times)

    def __array__(self, dtype=None, copy=None):
        return self.values if dtype is None else self.values.astype(dtype)


def dust_ode_control(max_steps=10000, atol=1e-6, rtol=1e-6, step_size_min=0,
                     step_size_max=np.inf, critical_times=None, debug_record_step_times=False,
                     save_history=False):
    return OdeControl([max_steps, atol, rtol, step_size_min, step_size_max], critical_times,
                      debug_record_step_times, save_history)


def _ode_control(ctl):
    if ctl is None:
        return dust_ode_control()
    return ctl if isinstance(ctl, OdeControl) else OdeControl(ctl)


class Unfilter:
    """dust_unfilter_create (R/interface-unfilter.R:20-71): the deterministic
    likelihood.  For the SIR ODE it integrates with Dormand-Prince 5(4); for
    the discrete-time SIR it runs the stochastic model with every sampler
    returning its mean, and can compute the adjoint gradient.  The
    likelihood of each particle is returned (P identical particles per group
    give identical values)."""

    def __init__(self, generator, time_start, data, n_particles, n_groups, dt, ode_control,
                 precision, shared_data=None, adjoint_form="reference"):
        if adjoint_form not in ("reference", "exact"):
            raise ValueError("'adjoint_form' must be 'reference' or 'exact'")
        # discrete SIR gradient: the reference's hand-written adjoint
        # (inst/examples/sir.cpp:126-198) or the exact derivative
        self.adjoint_form = adjoint_form
        if generator.name == "sirode":
            if dt is not None:
                raise ValueError("Can't use 'dt' with continuous-time systems")
        elif generator.name == "sir":
            if ode_control is not None:
                raise ValueError("Can't use 'ode_control' with discrete-time systems")
            if _precision(precision) != FP64:
                raise ValueError("The discrete unfilter runs in FP64 only")
        else:
            raise ValueError(f"No unfilter kernels for '{generator.name}'")
        self.generator = generator
        self.continuous = generator.name == "sirode"
        self.time_start = float(time_start)
        self.time, self.data, self.n_groups_data = prepare_data(data, n_groups,
                                                                shared_data=shared_data)
        check_time_start(time_start, self.time)
        self.n_particles = int(n_particles)
        self.n_groups = int(n_groups)
        self.dt = 1.0 if dt is None else float(dt)
        self.ode_control = _ode_control(ode_control)
        self.control = self.ode_control.values
        self.precision = _precision(precision)
        self.ptr = None
        self.rows = None
        self.last_index = None
        self.last_index_state = None
        self.last_snapshots = None

    @property
    def initialised(self):
        return self.ptr is not None

    def _prefix(self):
        return "dust2_unfilter_sirode" if self.continuous else "dust2_unfilter_sir"

    def _alloc(self, pars):
        self.rows = pars_matrix(self.generator, pars, self.n_groups)
        out = C.c_void_p()
        if self.continuous:
            call("dust2_unfilter_sirode_alloc", dptr(self.rows), self.n_groups,
                 self.n_particles, self.time_start, dptr(self.time), self.time.size,
                 dptr(self.data), self.n_groups_data, dptr(self.control), self.precision,
                 C.byref(out))
            ct = self.ode_control.critical_times
            if ct.size:
                call("dust2_unfilter_sirode_set_critical_times", out, dptr(ct), ct.size)
        else:
            call("dust2_unfilter_sir_alloc", dptr(self.rows), self.n_groups, self.n_particles,
                 self.time_start, dptr(self.time), self.time.size, self.dt, dptr(self.data),
                 self.n_groups_data, C.byref(out))
            call("dust2_unfilter_sir_set_adjoint_form", out,
                 int(self.adjoint_form == "exact"))
        self.ptr = out

    def run(self, pars, index_group=None, initial=None, adjoint=False,
            save_trajectories=False, save_snapshots=None, index_state=None):
        idx, n_idx, n_out = _index_group(index_group, self.n_groups)
        save = _SaveArgs(save_snapshots, index_state)
        if self.ptr is None:
            if pars is None:
                raise ValueError("'pars' cannot be NULL, as 'obj' is not initialised")
            self._alloc(pars)
        elif pars is not None:
            sel = np.arange(self.n_groups) if idx is None else idx.astype(np.int64)
            rows = pars_update(self.generator, pars, self.rows[sel])
            call(self._prefix() + "_update_pars", self.ptr, dptr(rows), szptr(idx), n_idx)
            self.rows[sel] = rows
        ll = np.zeros((n_out, self.n_particles))
        self.last_snapshots = None
        snaps, n_snaps, states, n_states = save.args
        if self.continuous:
            if initial is not None:
                raise ValueError("'initial' is not supported for the ODE unfilter")
            # adjoint = True: the gradient by forward sensitivities
            name = "dust2_unfilter_sirode_run_gradient" if adjoint else "dust2_unfilter_sirode_run"
            call(name, self.ptr, int(bool(save_trajectories)), snaps, n_snaps, states, n_states,
                 szptr(idx), n_idx, dptr(ll))
        else:
            init = None
            if initial is not None:
                init = np.ascontiguousarray(np.broadcast_to(
                    f64(initial), (self.n_groups, self.n_particles, 5)))
            call("dust2_unfilter_sir_run", self.ptr, dptr(init), int(bool(save_trajectories)),
                 snaps, n_snaps, int(bool(adjoint)), states, n_states, szptr(idx), n_idx,
                 dptr(ll))
        self.last_index = index_group
        self.last_index_state = save.states
        self.last_snapshots = save.snapshots
        if self.n_particles == 1:
            ll = ll[:, 0]
        return float(ll[0]) if (self.n_groups == 1 and ll.ndim == 1 and index_group is None) \
            else ll

    def last_gradient(self):
        """dust_likelihood_last_gradient (R/interface-likelihood.R): d ll /
        d (beta, gamma, I0) per group of the last run."""
        if self.ptr is None:
            raise ValueError("System was not run with 'adjoint = TRUE'")
        n = self.n_groups if self.last_index is None else np.size(self.last_index)
        if self.continuous:
            out = np.zeros((n, self.n_particles, 3))
            call("dust2_unfilter_sirode_last_gradient", self.ptr, dptr(out))
            out = out[:, 0] if self.n_particles == 1 else out
        else:
            out = np.zeros((n, 3))
            call("dust2_unfilter_sir_last_gradient", self.ptr, dptr(out))
        return out[0] if (self.n_groups == 1 and self.last_index is None) else out

    def last_state(self):
        out = np.zeros((self.n_groups, self.n_particles, 5))
        call(self._prefix() + "_last_state", self.ptr, dptr(out))
        return out

    def last_trajectories(self):
        """State at every data time of the last run (no reordering):
        (n_time, n_groups of the last run, n_particles, n_state)."""
        if self.ptr is None:
            raise DustError("Trajectories are not current")
        n = self.n_groups if self.last_index is None else np.size(self.last_index)
        out = np.zeros((self.time.size, n, self.n_particles, _traj_width(self)))
        call(self._prefix() + "_last_trajectories", self.ptr, dptr(out))
        return out

    def last_snapshots_state(self):
        """The whole state at the snapshot times of the last run:
        (n_snapshots, n_groups of the last run, n_particles, n_state)."""
        if self.ptr is None or self.last_snapshots is None:
            raise DustError("Snapshots are not current")
        n = self.n_groups if self.last_index is None else np.size(self.last_index)
        out = np.zeros((self.last_snapshots.size, n, self.n_particles, 5))
        call(self._prefix() + "_last_snapshots", self.ptr, dptr(out))
        return out

    def last_stats(self):
        out = np.zeros((self.n_groups * self.n_particles, 4))
        ms = np.zeros(1)
        call("dust2_unfilter_sirode_last_stats", self.ptr, dptr(out), dptr(ms))
        return out, float(ms[0])

    def __del__(self):
        if getattr(self, "ptr", None) is not None and _lib._lib is not None:
            getattr(_lib.lib(), self._prefix() + "_free")(self.ptr)
            self.ptr = None


def dust_unfilter_create(generator, time_start, data, n_particles=1, n_groups=1, dt=None,
                         ode_control=None, precision="fp64", shared_data=None,
                         adjoint_form="reference"):
    """R/interface-unfilter.R:20-71; shared_data as dust_filter_create.
    adjoint_form (discrete SIR): 'reference' for the reference's hand-written
    adjoint gradient, 'exact' for the derivative of the likelihood."""
    if generator.name in ModelFilter.DATA_FIELDS:
        return ModelUnfilter(generator, time_start, data, n_particles, n_groups, dt, ode_control,
                             shared_data)
    return Unfilter(generator, time_start, data, n_particles, n_groups, dt, ode_control,
                    precision, shared_data, adjoint_form=adjoint_form)


def dust_likelihood_last_gradient(obj):
    return obj.last_gradient()


# ---------------------------------------------------------------------------
# Kernel-level helpers used by the parity tests
# ---------------------------------------------------------------------------
def gpu_resample(log_weights, u):
    w = f64(log_weights).reshape(-1)
    idx = np.zeros(w.size, dtype=np.uint32)
    out = np.zeros(3)
    call("dust2_amd_resample", dptr(w), w.size, float(u), u32ptr(idx), dptr(out))
    return idx, out


def gpu_trajectory_lineage(raw, ancestry=None, select=None):
    """The lineage-export kernel on given states: raw (T, G, P, S),
    ancestry (T, G, P) or None, select (G,) or None."""
    raw = f64(raw)
    T, G, P, S = raw.shape
    anc = None if ancestry is None else np.ascontiguousarray(ancestry, dtype=np.uint32)
    sel = None if select is None else np.ascontiguousarray(select, dtype=np.uintp)
    out = np.zeros((T, G, 1 if sel is not None else P, S))
    call("dust2_amd_trajectory_lineage", dptr(raw), u32ptr(anc), T, G, P, S, szptr(sel),
         dptr(out))
    return out


INTERP_KINDS = {"constant": 0, "linear": 1, "spline": 2}


def gpu_interpolate(kind, t, y, z):
    """Evaluate a piecewise constant / linear / natural cubic spline
    interpolator on the device (dust2/interpolate/interpolate.hpp): t
    (n_t,) strictly increasing, y (n_t,) or (n_t, n_series), z (n_z,).
    Returns (n_z,) or (n_z, n_series)."""
    t = f64(t).reshape(-1)
    y = f64(y)
    series = y.ndim > 1
    y2 = y.reshape(y.shape[0], -1) if y.ndim else y.reshape(1, 1)
    n_series = y2.shape[1]
    yv = np.ascontiguousarray(y2).reshape(-1)
    z = f64(z).reshape(-1)
    out = np.zeros(max(z.size * n_series, 1))
    call("dust2_amd_interpolate", INTERP_KINDS[kind], dptr(t), t.size, dptr(yv), yv.size,
         n_series, dptr(z), z.size, dptr(out))
    out = out[:z.size * n_series].reshape(z.size, n_series)
    return out if series else out[:, 0]


def rng_jump(state, long_jump=False):
    """monty's jump() (2^128 draws) or long_jump() (2^192 draws) applied to
    each 4-word xoshiro256+ state (R/monty.R:29-36)."""
    st = np.array(state, dtype=np.uint64).reshape(-1, 4).copy()
    call("dust2_amd_rng_jump", u64ptr(st), st.shape[0], int(bool(long_jump)))
    return st


def gpu_rng_streams(n_streams, seed=None, seed_state=None):
    out = np.zeros(n_streams * 4, dtype=np.uint64)
    raw = None if seed_state is None else np.ascontiguousarray(seed_state, dtype=np.uint64)
    n_raw = 0 if raw is None else raw.size // 4
    call("dust2_amd_rng_streams", u64ptr(raw), n_raw, C.c_uint64(0 if seed is None else seed),
         n_streams, u64ptr(out))
    return out.reshape(n_streams, 4)


SAMPLER_KINDS = {"binomial": 0, "exponential": 1, "normal": 2, "poisson": 3, "real": 4}


def gpu_sample(kind, rng, n_draws, a, b=None, precision="fp64", deterministic=False):
    """Draw n_draws values per stream; rng (n_streams x 4) is advanced in place."""
    rng = np.ascontiguousarray(rng, dtype=np.uint64)
    n = rng.size // 4
    a = np.broadcast_to(f64(a), (n,)).copy()
    b = np.broadcast_to(f64(0.0 if b is None else b), (n,)).copy()
    out = np.zeros((n, n_draws))
    err = np.zeros(n, dtype=np.uint32)
    call("dust2_amd_sample", SAMPLER_KINDS[kind], _precision(precision), u64ptr(rng), n,
         n_draws, dptr(a), dptr(b), int(bool(deterministic)), dptr(out), u32ptr(err))
    return out, err, rng
