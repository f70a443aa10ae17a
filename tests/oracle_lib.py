"""ctypes access to the CPU oracle (oracle/liboracle.so) and, where it was
built, to the reference's own compiled parts (oracle/_ref/libdust2ref.so).

Test infrastructure only: the product never imports this module.
"""
import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_SO = os.path.join(ORACLE_DIR, "liboracle.so")
REF_SO = os.path.join(ORACLE_DIR, "_ref", "libdust2ref.so")

D = C.POINTER(C.c_double)
U64 = C.POINTER(C.c_uint64)
U32 = C.POINTER(C.c_uint32)
SZ = C.POINTER(C.c_size_t)

_oracle = None
_ref = None


def _dp(a):
    return None if a is None else a.ctypes.data_as(D)


def _u64(a):
    return None if a is None else a.ctypes.data_as(U64)


def _u32(a):
    return None if a is None else a.ctypes.data_as(U32)


def build_oracle():
    subprocess.run(["make", "-s", "-C", ORACLE_DIR, "liboracle.so"], check=True)


def oracle():
    global _oracle
    if _oracle is None:
        if not os.path.exists(ORACLE_SO):
            build_oracle()
        _oracle = C.CDLL(ORACLE_SO)
        o = _oracle
        o.oracle_dpois_log.restype = C.c_double
        o.oracle_dpois_log.argtypes = [C.c_double, C.c_double]
        o.oracle_scale_log_weights.restype = C.c_double
        o.oracle_scale_log_weights.argtypes = [C.c_size_t, D]
        o.oracle_resample_weight.argtypes = [C.c_size_t, D, C.c_double, U32]
        o.oracle_sirode_unfilter_gradient.argtypes = [D, C.c_size_t, C.c_double, D, C.c_size_t,
                                                      D, C.c_size_t, D, C.c_int, D, D,
                                                      C.c_char_p, C.c_size_t]
        o.oracle_trajectory_lineage.argtypes = [D, U32, C.c_size_t, C.c_size_t, C.c_size_t,
                                                C.c_size_t, SZ, D]
        o.oracle_trajectory_lineage.restype = None
        o.oracle_interpolate.argtypes = [C.c_int, D, C.c_size_t, D, C.c_size_t, D, C.c_size_t,
                                         D, SZ]
        o.oracle_seed_words.argtypes = [C.c_uint64, U64]
        o.oracle_prng.argtypes = [U64, C.c_size_t, C.c_size_t, U64]
        o.oracle_jump.argtypes = [U64, C.c_int]
        o.oracle_rng_next.argtypes = [U64, C.c_size_t, U64]
        o.oracle_random_real.argtypes = [U64, C.c_size_t, D]
        o.oracle_sample.argtypes = [C.c_int, U64, C.c_int, C.c_size_t, D, D, D, C.c_char_p,
                                    C.c_size_t]
        filt = [D, C.c_size_t, C.c_size_t, C.c_double, D, C.c_size_t, C.c_double, D,
                C.c_size_t, U64, C.c_uint64, C.c_int, D, SZ, C.c_size_t, D, D, U64, U32, D,
                C.c_char_p, C.c_size_t]
        o.oracle_sir_filter_run.argtypes = filt
        o.oracle_seir16_filter_run.argtypes = filt
        o.oracle_sir_tv_filter_run.argtypes = filt
        sim = [D, C.c_size_t, C.c_size_t, C.c_double, C.c_double, D, C.c_size_t, U64, C.c_int,
               C.c_int, D, U64, C.c_char_p, C.c_size_t]
        o.oracle_sir_simulate.argtypes = sim
        o.oracle_seir16_simulate.argtypes = sim
        o.oracle_sir_tv_simulate.argtypes = sim
        o.oracle_walk_simulate.argtypes = [C.c_size_t] + sim
        unf = [D, C.c_size_t, C.c_size_t, C.c_double, D, C.c_size_t, C.c_double, D,
               C.c_size_t, C.c_int, D, D, D, C.c_char_p, C.c_size_t]
        o.oracle_sir_unfilter_det_run.argtypes = unf
        o.oracle_seir16_unfilter_run.argtypes = unf
        o.oracle_sirode_unfilter_run.argtypes = [D, C.c_size_t, C.c_size_t, C.c_double, D,
                                                 C.c_size_t, D, C.c_size_t, D, C.c_int, D, D,
                                                 D, C.c_char_p, C.c_size_t]
        o.oracle_sir_unfilter_run.argtypes = [D, C.c_size_t, C.c_double, D, C.c_size_t,
                                              C.c_double, D, D, D, C.c_int, C.c_char_p,
                                              C.c_size_t]
    return _oracle


def ref_available():
    return os.path.exists(REF_SO)


def ref():
    global _ref
    if _ref is None:
        _ref = C.CDLL(REF_SO)
        _ref.ref_scale_log_weights.restype = C.c_double
        _ref.ref_scale_log_weights.argtypes = [C.c_size_t, D]
        _ref.ref_resample_weight.argtypes = [C.c_size_t, D, C.c_double, U32]
        _ref.ref_sirode_unfilter.argtypes = [D, C.c_size_t, C.c_double, D, C.c_size_t, D, D, D,
                                             D, D, C.c_char_p, C.c_size_t]
        _ref.ref_logistic.argtypes = [C.c_double, C.c_double, C.c_double, C.c_double, D,
                                      C.c_size_t, D, D]
        _ref.ref_sirode_system.argtypes = [D, C.c_double, D, C.c_size_t, D, D, C.c_size_t,
                                           C.c_size_t, D, D, D, C.c_size_t, SZ, C.c_char_p,
                                           C.c_size_t]
        _ref.ref_interpolate.argtypes = [C.c_int, D, C.c_size_t, D, C.c_size_t, D, C.c_size_t,
                                         D, C.c_char_p, C.c_size_t]
    return _ref


class OracleError(RuntimeError):
    pass


def _check(rc, buf):
    if rc != 0:
        raise OracleError(buf.value.decode())


# ---- thin numpy wrappers ---------------------------------------------------
def seed_words(seed):
    out = np.zeros(4, dtype=np.uint64)
    oracle().oracle_seed_words(C.c_uint64(seed), _u64(out))
    return out


def prng(n, seed=None, seed_state=None):
    s = seed_words(seed) if seed_state is None else np.ascontiguousarray(seed_state,
                                                                       dtype=np.uint64)
    out = np.zeros(n * 4, dtype=np.uint64)
    oracle().oracle_prng(_u64(s), s.size // 4, n, _u64(out))
    return out.reshape(n, 4)


def jump(state, long=False):
    s = np.array(state, dtype=np.uint64).copy()
    oracle().oracle_jump(_u64(s), int(long))
    return s


def rng_next(state, n):
    s = np.array(state, dtype=np.uint64).copy()
    out = np.zeros(n, dtype=np.uint64)
    oracle().oracle_rng_next(_u64(s), n, _u64(out))
    return out, s


def random_real(state, n):
    s = np.array(state, dtype=np.uint64).copy()
    out = np.zeros(n)
    oracle().oracle_random_real(_u64(s), n, _dp(out))
    return out, s


KINDS = {"binomial": 0, "exponential": 1, "normal": 2, "poisson": 3}


def sample(kind, state, a, b=None, deterministic=False):
    """Draw len(a) values from ONE stream with per-draw arguments."""
    s = np.array(state, dtype=np.uint64).copy()
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.zeros_like(a) if b is None else np.ascontiguousarray(
        np.broadcast_to(b, a.shape), dtype=np.float64)
    out = np.zeros(a.size)
    buf = C.create_string_buffer(512)
    rc = oracle().oracle_sample(KINDS[kind], _u64(s), int(deterministic), a.size, _dp(a), _dp(b),
                                _dp(out), buf, 512)
    _check(rc, buf)
    return out, s


def dpois_log(x, lam):
    return oracle().oracle_dpois_log(float(x), float(lam))


def scale_log_weights(w):
    w = np.array(w, dtype=np.float64)
    r = oracle().oracle_scale_log_weights(w.size, _dp(w))
    return r, w


def resample_weight(w, u):
    w = np.ascontiguousarray(w, dtype=np.float64)
    idx = np.zeros(w.size, dtype=np.uint32)
    oracle().oracle_resample_weight(w.size, _dp(w), float(u), _u32(idx))
    return idx


def sir_filter(pars, n_particles, time_start, time, dt, data, rng=None, seed=42, n_threads=1,
               initial=None, index_group=None, ancestry=False, model="sir", traj=False):
    """pars: (G, n_pars); data: (n_time, G_data) (or (n_time, G_data, n_data)).
    Returns dict(ll, state (G,P,S), rng (G*(P+2),4), ancestry (T,G,P) and
    traj (T,G,P,S), the states after each time's resampling, when asked)."""
    pars = np.ascontiguousarray(pars, dtype=np.float64)
    G = pars.shape[0]
    P = n_particles
    time = np.ascontiguousarray(time, dtype=np.float64)
    data = np.ascontiguousarray(data, dtype=np.float64)
    n_groups_data = data.shape[1]
    S = 80 if model == "seir16" else 5
    ll = np.zeros(G)
    state = np.zeros((G, P, S))
    rng_out = np.zeros((G * (P + 2), 4), dtype=np.uint64)
    anc = np.zeros((time.size, G, P), dtype=np.uint32) if ancestry else None
    raw = np.zeros((time.size, G, P, S)) if traj else None
    rng_in = None if rng is None else np.ascontiguousarray(rng, dtype=np.uint64)
    init = None if initial is None else np.ascontiguousarray(initial, dtype=np.float64)
    idx = None if index_group is None else np.ascontiguousarray(index_group, dtype=np.uintp)
    buf = C.create_string_buffer(4096)
    fn = {"sir": oracle().oracle_sir_filter_run, "seir16": oracle().oracle_seir16_filter_run,
          "sir_tv": oracle().oracle_sir_tv_filter_run}[model]
    rc = fn(_dp(pars), G, P, float(time_start), _dp(time), time.size, float(dt), _dp(data),
            n_groups_data, _u64(rng_in), C.c_uint64(seed), n_threads, _dp(init),
            None if idx is None else idx.ctypes.data_as(SZ), 0 if idx is None else idx.size,
            _dp(ll), _dp(state), _u64(rng_out), _u32(anc), _dp(raw), buf, 4096)
    _check(rc, buf)
    return {"ll": ll, "state": state, "rng": rng_out, "ancestry": anc, "traj": raw}


def sir_simulate(pars, n_particles, time_start, dt, times, rng, deterministic=False,
                 n_threads=1, model="sir"):
    pars = np.ascontiguousarray(pars, dtype=np.float64)
    G = pars.shape[0]
    P = n_particles
    S = 80 if model == "seir16" else 5
    times = np.ascontiguousarray(times, dtype=np.float64)
    rng = np.ascontiguousarray(rng, dtype=np.uint64)
    out = np.zeros((times.size, G, P, S))
    rng_out = np.zeros((G * P, 4), dtype=np.uint64)
    buf = C.create_string_buffer(4096)
    fn = {"sir": oracle().oracle_sir_simulate, "seir16": oracle().oracle_seir16_simulate,
          "sir_tv": oracle().oracle_sir_tv_simulate}[model]
    rc = fn(_dp(pars), G, P, float(time_start), float(dt), _dp(times), times.size, _u64(rng),
            int(deterministic), n_threads, _dp(out), _u64(rng_out), buf, 4096)
    _check(rc, buf)
    return out, rng_out


def walk_simulate(pars, n_particles, time_start, dt, times, rng, deterministic=False):
    """The random walk as a system; pars (G, 3): sd, len, random_initial
    (len the same for every group).  Returns (n_times, G, P, len) and the
    streams after."""
    pars = np.ascontiguousarray(pars, dtype=np.float64)
    G = pars.shape[0]
    L = int(pars[0, 1])
    times = np.ascontiguousarray(times, dtype=np.float64)
    rng = np.ascontiguousarray(rng, dtype=np.uint64)
    out = np.zeros((times.size, G, n_particles, L))
    rng_out = np.zeros((G * n_particles, 4), dtype=np.uint64)
    buf = C.create_string_buffer(4096)
    rc = oracle().oracle_walk_simulate(L, _dp(pars), G, n_particles, float(time_start),
                                       float(dt), _dp(times), times.size, _u64(rng),
                                       int(deterministic), 1, _dp(out), _u64(rng_out), buf,
                                       4096)
    _check(rc, buf)
    return out, rng_out


def discrete_unfilter(pars, n_particles, time_start, time, dt, data, model="sir",
                      n_threads=1, trajectories=False):
    """Deterministic unfilter of a discrete-time model: per-particle ll (G*P),
    final state (G, P, S) and, when asked, the state at every data time
    (n_time, G, P, S).  data: (n_time, G_data[, n_data])."""
    pars = np.ascontiguousarray(pars, dtype=np.float64)
    G = pars.shape[0]
    P = n_particles
    S = 80 if model == "seir16" else 5
    time = np.ascontiguousarray(time, dtype=np.float64)
    data = np.ascontiguousarray(data, dtype=np.float64)
    n_groups_data = data.shape[1]
    ll = np.zeros(G * P)
    state = np.zeros((G, P, S))
    traj = np.zeros((time.size, G, P, S)) if trajectories else None
    buf = C.create_string_buffer(4096)
    fn = {"sir": oracle().oracle_sir_unfilter_det_run,
          "seir16": oracle().oracle_seir16_unfilter_run}[model]
    rc = fn(_dp(pars), G, P, float(time_start), _dp(time), time.size, float(dt), _dp(data),
            n_groups_data, n_threads, _dp(ll), _dp(state), _dp(traj), buf, 4096)
    _check(rc, buf)
    return {"ll": ll, "state": state, "traj": traj}


def sirode_unfilter(pars, n_particles, time_start, time, data, control=None, n_threads=1):
    pars = np.ascontiguousarray(pars, dtype=np.float64)
    G = pars.shape[0]
    P = n_particles
    time = np.ascontiguousarray(time, dtype=np.float64)
    data = np.ascontiguousarray(data, dtype=np.float64)
    n_groups_data = 1 if data.ndim == 1 else data.shape[1]
    ctl = np.array([10000, 1e-6, 1e-6, 0, np.inf] if control is None else control,
                   dtype=np.float64)
    ll = np.zeros(G * P)
    state = np.zeros((G, P, 5))
    stats = np.zeros((G * P, 4))
    buf = C.create_string_buffer(4096)
    rc = oracle().oracle_sirode_unfilter_run(_dp(pars), G, P, float(time_start), _dp(time),
                                             time.size, _dp(data), n_groups_data, _dp(ctl),
                                             n_threads, _dp(ll), _dp(state), _dp(stats), buf,
                                             4096)
    _check(rc, buf)
    return {"ll": ll, "state": state, "stats": stats}


def trajectory_lineage(raw, ancestry=None, select=None):
    """raw (T, G, P, S); ancestry (T, G, P) or None; select (G,) or None."""
    raw = np.ascontiguousarray(raw, dtype=np.float64)
    T, G, P, S = raw.shape
    anc = None if ancestry is None else np.ascontiguousarray(ancestry, dtype=np.uint32)
    sel = None if select is None else np.ascontiguousarray(select, dtype=np.uintp)
    out = np.zeros((T, G, 1 if sel is not None else P, S))
    oracle().oracle_trajectory_lineage(_dp(raw), _u32(anc), T, G, P, S,
                                       None if sel is None else sel.ctypes.data_as(SZ),
                                       _dp(out))
    return out


def sirode_gradient(pars, time_start, time, data, control=None, n_threads=1):
    """Log-likelihood and d ll / d (beta, gamma, I0) of the SIR ODE
    unfilter by forward sensitivities; pars (G, 4), data (n_time,) or
    (n_time, G)."""
    pars = np.ascontiguousarray(pars, dtype=np.float64)
    G = pars.shape[0]
    time = np.ascontiguousarray(time, dtype=np.float64)
    data = np.ascontiguousarray(data, dtype=np.float64)
    n_groups_data = 1 if data.ndim == 1 else data.shape[1]
    ctl = None if control is None else np.ascontiguousarray(control, dtype=np.float64)
    ll = np.zeros(G)
    grad = np.zeros((G, 3))
    buf = C.create_string_buffer(4096)
    rc = oracle().oracle_sirode_unfilter_gradient(_dp(pars), G, float(time_start), _dp(time),
                                                  time.size, _dp(data), n_groups_data,
                                                  _dp(ctl), n_threads, _dp(ll), _dp(grad),
                                                  buf, 4096)
    _check(rc, buf)
    return ll, grad


def sir_unfilter(pars, time_start, time, dt, data, gradient=True, exact=False):
    """Deterministic SIR unfilter and its adjoint gradient (beta, gamma, I0)
    per group: the reference's hand-written adjoint (exact=False) or the
    derivative of the likelihood (exact=True)."""
    pars = np.ascontiguousarray(pars, dtype=np.float64)
    G = pars.shape[0]
    time = np.ascontiguousarray(time, dtype=np.float64)
    data = np.ascontiguousarray(np.broadcast_to(
        np.asarray(data, dtype=np.float64).reshape(time.size, -1), (time.size, G)))
    ll = np.zeros(G)
    grad = np.zeros((G, 3)) if gradient else None
    buf = C.create_string_buffer(4096)
    rc = oracle().oracle_sir_unfilter_run(_dp(pars), G, float(time_start), _dp(time),
                                          time.size, float(dt), _dp(data), _dp(ll), _dp(grad),
                                          int(bool(exact)), buf, 4096)
    _check(rc, buf)
    return ll, grad


# ---- reference (compiled from /root/reference headers) ---------------------
def ref_scale_log_weights(w):
    w = np.array(w, dtype=np.float64)
    r = ref().ref_scale_log_weights(w.size, _dp(w))
    return r, w


def ref_resample_weight(w, u):
    w = np.ascontiguousarray(w, dtype=np.float64)
    idx = np.zeros(w.size, dtype=np.uint32)
    ref().ref_resample_weight(w.size, _dp(w), float(u), _u32(idx))
    return idx


def ref_sirode(pars, time_start, time, data, control=None):
    pars = np.ascontiguousarray(pars, dtype=np.float64)
    n = pars.shape[0]
    time = np.ascontiguousarray(time, dtype=np.float64)
    data = np.ascontiguousarray(data, dtype=np.float64)
    ctl = np.array([10000, 1e-6, 1e-6, 0, np.inf] if control is None else control,
                   dtype=np.float64)
    ll = np.zeros(n)
    state = np.zeros((n, time.size, 5))
    stats = np.zeros((n, 4))
    buf = C.create_string_buffer(512)
    rc = ref().ref_sirode_unfilter(_dp(pars), n, float(time_start), _dp(time), time.size,
                                   _dp(data), _dp(ctl), _dp(ll), _dp(state), _dp(stats), buf, 512)
    _check(rc, buf)
    return {"ll": ll, "state": state, "stats": stats}


def ref_logistic(r1, r2, K, y0, time, control=None):
    time = np.ascontiguousarray(time, dtype=np.float64)
    ctl = np.array([10000, 1e-6, 1e-6, 0, np.inf] if control is None else control,
                   dtype=np.float64)
    out = np.zeros((time.size, 2))
    ref().ref_logistic(r1, r2, K, y0, _dp(time), time.size, _dp(ctl), _dp(out))
    return out


INTERP_KINDS = {"constant": 0, "linear": 1, "spline": 2}


def _interp_args(t, y, z):
    t = np.ascontiguousarray(t, dtype=np.float64).reshape(-1)
    y = np.asarray(y, dtype=np.float64)
    n_series = 1 if y.ndim == 1 else y.shape[1]
    yv = np.ascontiguousarray(y.reshape(-1))
    z = np.ascontiguousarray(z, dtype=np.float64).reshape(-1)
    return t, yv, n_series, z, y.ndim > 1


def interpolate(kind, t, y, z):
    """Oracle interpolation; raises OracleError naming the first time
    outside the knots ("before" / "after")."""
    t, yv, n_series, z, series = _interp_args(t, y, z)
    out = np.zeros(max(z.size * n_series, 1))
    bad = np.zeros(1, dtype=np.uintp)
    rc = oracle().oracle_interpolate(INTERP_KINDS[kind], _dp(t), t.size, _dp(yv), n_series,
                                     _dp(z), z.size, _dp(out), bad.ctypes.data_as(SZ))
    if rc:
        raise OracleError(("before" if rc == 1 else "after") + f" at {z[bad[0]]}")
    out = out[:z.size * n_series].reshape(z.size, n_series)
    return out if series else out[:, 0]


def ref_interpolate(kind, t, y, z):
    """The reference's own interpolators (oracle/_ref)."""
    t, yv, n_series, z, series = _interp_args(t, y, z)
    out = np.zeros(max(z.size * n_series, 1))
    buf = C.create_string_buffer(512)
    rc = ref().ref_interpolate(INTERP_KINDS[kind], _dp(t), t.size, _dp(yv), n_series, _dp(z),
                               z.size, _dp(out), buf, 512)
    _check(rc, buf)
    out = out[:z.size * n_series].reshape(z.size, n_series)
    return out if series else out[:, 0]


def sir_tv_rows(pars, kind, beta_t, beta_y):
    """Oracle parameter rows of the SIR with interpolated beta: pars (G, 5)
    SIR rows (beta unused), beta_t (n,), beta_y (G, n) or (n,)."""
    kinds = {"constant": 0, "linear": 1, "spline": 2}
    pars = np.atleast_2d(np.asarray(pars, dtype=np.float64))
    beta_y = np.atleast_2d(np.asarray(beta_y, dtype=np.float64))
    G, n = pars.shape[0], len(beta_t)
    rows = np.zeros((G, 7 + 2 * 64))
    rows[:, :5] = pars
    rows[:, 5] = kinds[kind]
    rows[:, 6] = n
    rows[:, 7:7 + n] = beta_t
    rows[:, 7 + 64:7 + 64 + n] = np.broadcast_to(beta_y, (G, n))
    return rows


def ref_sirode_system(pars, time_start, targets, control=None, critical_times=(),
                      restart_after=None):
    """One SIRODE particle through the reference solver driven as a
    continuous-time system (oracle/_ref): states at the targets, the
    solver counters and the recorded step times."""
    pars = np.ascontiguousarray(pars, dtype=np.float64)
    targets = np.ascontiguousarray(targets, dtype=np.float64).reshape(-1)
    ctl = np.array([10000, 1e-6, 1e-6, 0, np.inf] if control is None else control,
                   dtype=np.float64)
    crit = np.ascontiguousarray(critical_times, dtype=np.float64).reshape(-1)
    state = np.zeros((targets.size, 5))
    stats = np.zeros(4)
    cap = 200000
    times = np.zeros(cap)
    n = np.zeros(1, dtype=np.uintp)
    buf = C.create_string_buffer(512)
    rc = ref().ref_sirode_system(_dp(pars), float(time_start), _dp(targets), targets.size,
                                 _dp(ctl), _dp(crit) if crit.size else None, crit.size,
                                 C.c_size_t(targets.size if restart_after is None
                                            else restart_after),
                                 _dp(state), _dp(stats), _dp(times), cap,
                                 n.ctypes.data_as(SZ), buf, 512)
    _check(rc, buf)
    return {"state": state, "stats": stats, "step_times": times[:int(n[0])].copy()}


GENERAL_IDS = {"bounce": 0, "change": 1, "eventzero": 2, "specialevent": 3, "sirdelay": 4,
               "logdelay": 5}


def ref_ode_general(model, row, targets, y0=None, control=None, critical_times=(),
                    save_history=False, time_start=0.0):
    """One particle of a continuous-time model with events, special
    variables, resets or delays through the reference solver (oracle/_ref):
    states at the targets, counters, step times, events (time, 0-based
    index, sign) and the saved history (t0, t1, h, coefficients)."""
    row = np.ascontiguousarray(row, dtype=np.float64)
    targets = np.ascontiguousarray(targets, dtype=np.float64).reshape(-1)
    ctl = np.array([10000, 1e-6, 1e-6, 0, np.inf] if control is None else control,
                   dtype=np.float64)
    crit = np.ascontiguousarray(critical_times, dtype=np.float64).reshape(-1)
    S = {"bounce": 2, "change": 1, "eventzero": 2, "specialevent": 3, "sirdelay": 5,
         "logdelay": 1}[model]
    H = {"bounce": 2, "change": 1, "eventzero": 2, "specialevent": 3, "sirdelay": 1,
         "logdelay": 1}[model]
    y0a = None if y0 is None else np.ascontiguousarray(y0, dtype=np.float64)
    state = np.zeros((targets.size, S))
    stats = np.zeros(4)
    cap, cap_ev, cap_h = 200000, 10000, 200000
    times = np.zeros(cap)
    ev = np.zeros(cap_ev * 3)
    hist = np.zeros(cap_h * (3 + 5 * H))
    n = np.zeros(3, dtype=np.uintp)
    buf = C.create_string_buffer(512)
    szp = [n[i:i + 1].ctypes.data_as(SZ) for i in range(3)]
    rc = ref().ref_ode_general(
        GENERAL_IDS[model], _dp(row), _dp(y0a) if y0a is not None else None,
        C.c_double(time_start), _dp(targets), C.c_size_t(targets.size), _dp(ctl),
        _dp(crit) if crit.size else None, C.c_size_t(crit.size), C.c_int(int(save_history)),
        _dp(state), _dp(stats), _dp(times), C.c_size_t(cap), szp[0], _dp(ev),
        C.c_size_t(cap_ev), szp[1], _dp(hist), C.c_size_t(cap_h), szp[2], buf, 512)
    _check(rc, buf)
    n_ev, n_h = int(n[1]), int(n[2])
    return {"state": state, "stats": stats, "step_times": times[:int(n[0])].copy(),
            "events": ev[:3 * n_ev].reshape(n_ev, 3).copy(),
            "history": hist[:n_h * (3 + 5 * H)].reshape(n_h, 3 + 5 * H).copy()}


def ref_ode_mixed(model, row, targets, rng, y0=None, dt=1.0, deterministic=False,
                  time_start=0.0):
    """One particle of a mixed-time model (special, malaria) through the
    reference solver with the mixed-time loop of the continuous system
    restated (oracle/_ref); rng: the particle's stream (4 words)."""
    ids = {"special": 6, "malaria": 7}
    S = {"special": 2, "malaria": 20}[model]
    row = np.ascontiguousarray(row, dtype=np.float64)
    targets = np.ascontiguousarray(targets, dtype=np.float64).reshape(-1)
    r = np.ascontiguousarray(rng, dtype=np.uint64).reshape(4)
    y0a = None if y0 is None else np.ascontiguousarray(y0, dtype=np.float64)
    state = np.zeros((targets.size, S))
    stats = np.zeros(4)
    cap = 200000
    times = np.zeros(cap)
    n = np.zeros(1, dtype=np.uintp)
    buf = C.create_string_buffer(512)
    rc = ref().ref_ode_mixed(
        C.c_int(ids[model]), _dp(row), _dp(y0a) if y0a is not None else None,
        C.c_double(time_start), C.c_double(dt), _dp(targets), C.c_size_t(targets.size),
        _u64(r), C.c_int(int(deterministic)), _dp(state), _dp(stats), _dp(times),
        C.c_size_t(cap), n.ctypes.data_as(SZ), buf, 512)
    _check(rc, buf)
    return {"state": state, "stats": stats, "step_times": times[:int(n[0])].copy()}


def ref_malaria_filter(row, P, time_start, times, data, dt, part_rng, filt_rng):
    """One group of the malaria particle filter through the reference solver
    and the reference's own scale_log_weights / resample_weight
    (oracle/_ref): the log-likelihood, the final states [P][20] and the
    filter stream after the run."""
    row = np.ascontiguousarray(row, dtype=np.float64)
    times = np.ascontiguousarray(times, dtype=np.float64).reshape(-1)
    data = np.ascontiguousarray(data, dtype=np.float64).reshape(times.size, 2)
    pr = np.ascontiguousarray(part_rng, dtype=np.uint64).reshape(P, 4)
    fr = np.ascontiguousarray(filt_rng, dtype=np.uint64).reshape(4).copy()
    ll = np.zeros(1)
    state = np.zeros((P, 20))
    buf = C.create_string_buffer(512)
    rc = ref().ref_malaria_filter(_dp(row), C.c_size_t(P), C.c_double(time_start), _dp(times),
                                  C.c_size_t(times.size), _dp(data), C.c_double(dt), _u64(pr),
                                  _u64(fr), _dp(ll), _dp(state), buf, 512)
    _check(rc, buf)
    return float(ll[0]), state, fr

