"""Benchmarks of the batched particle engine on one MI355X.

Default workload (the headline): 512 parameter groups x 2048 particles of
the stochastic SIR model (beta ~ U(0.1, 0.5), gamma ~ U(0.05, 0.2),
N ~ U(1e3, 1e6) integer, I0 = 10), dt = 0.25, 365 daily Poisson
observations per group simulated from the model with a fixed seed, FP32
state and samplers with FP64 log-weights, resampling sums and likelihood.
One step = one complete filter run of every group (1.53e9 particle-steps).

Other workloads (--workload):
  sir-replicates  1000 replicate filters of the 1000-particle SIR problem
                  (100 daily observations), FP64: evals/s and the mean and
                  variance of the log-likelihood beside the FP64 CPU oracle
  sirode          65,536 SIR ODE parameter sets, Dormand-Prince 5(4),
                  rtol = atol = 1e-6, 365 daily observations, FP64 and FP32
  seir16          128 groups x 8192 particles of the 16-age SEIR (80
                  states), 200 weekly observations, FP32, batched vs one
                  group per launch
  serving         64 random-walk MCMC chains submitting likelihood
                  evaluations (sir / seir16 / sirode, particles in
                  {128, 512, 2048, 8192}, data length U{50..730}) at up to
                  2000 requests/s: continuous batching vs one request at a
                  time, evals/s and p50/p99 latency
  valu-peak       the vector-ALU issue rate of independent fused
                  multiply-adds (FP32, FP64) at 1-8 wavefronts per SIMD,
                  the compute ceiling of the sampler kernels

Each prints one JSON line.  `roofline` refers to the dominant kernel:
algorithmic bytes per launch (state gather-read and write, xoshiro256+
read and write, FP64 weight write and read, u32 ancestor write and read per
particle per observation) over its event-timed duration.
"""
import argparse
import json
import os
import sys
import time as _time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
SIR_NAMES = ["N", "I0", "beta", "gamma", "exp_noise"]


def cpu_threads():
    """nproc: the CPUs this process may run on."""
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def sized_sample(run, g_max, target_s):
    """The number of groups g whose oracle run takes about target_s: doubles
    g until a run takes 0.3 s, then scales."""
    g = 1
    while True:
        t0 = _time.perf_counter()
        run(g)
        el = _time.perf_counter() - t0
        if el > 0.3 or g >= g_max:
            break
        g = min(g_max, g * 2)
    return int(max(1, min(g_max, round(g * target_s / max(el, 1e-3)))))


def median_runs(run, g, n_runs):
    """Median wall time of n_runs calls run(g, k) (k = the replicate)."""
    times, outs = [], []
    for k in range(n_runs):
        t0 = _time.perf_counter()
        outs.append(run(g, k))
        times.append(_time.perf_counter() - t0)
    return float(np.median(times)), outs


def cpu_protocol(run, g_max, per_g, unit, n_runs, what, target_s=3.0, target_1_s=2.0):
    """The oracle's rate at nproc threads and at one thread: run(g, k, n_threads)
    does g work units (per_g units of `unit` each) with replicate k.  The
    sizing runs are the warm-up; then the median of n_runs runs."""
    threads = cpu_threads()
    g = sized_sample(lambda n: run(n, 0, threads), g_max, target_s)
    el, outs = median_runs(lambda n, k: run(n, k, threads), g, n_runs)
    g1 = sized_sample(lambda n: run(n, 0, 1), g_max, target_1_s)
    el1, _ = median_runs(lambda n, k: run(n, k, 1), g1, n_runs)
    return {"value": g * per_g / el, "unit": unit, "cores": threads, "kind": "port",
            "sample": f"{what(g)}, FP64 oracle (OpenMP, {threads} threads = nproc), "
                      f"median of {n_runs} runs of {el:.2f} s",
            "one_thread": g1 * per_g / el1,
            "one_thread_sample": f"{what(g1)}, median of {n_runs} runs of {el1:.2f} s",
            "cpu_model": cpu_model(),
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS")}, g, outs


def oracle():
    """The CPU oracle, used only for the cpu_baseline leg."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib
    return oracle_lib


# ---------------------------------------------------------------------------
# workload construction
# ---------------------------------------------------------------------------
def sir_groups(n_groups, seed=42):
    rng = np.random.default_rng(seed)
    beta = rng.uniform(0.1, 0.5, n_groups)
    gamma = rng.uniform(0.05, 0.2, n_groups)
    N = np.round(rng.uniform(1e3, 1e6, n_groups))
    return np.column_stack([N, np.full(n_groups, 10.0), beta, gamma, np.full(n_groups, 1e6)])


def simulate_sir_data(lib, pars, time, dt, seed):
    """Daily incidence from one stochastic run of each group's own model
    (GPU system kernels)."""
    plist = [dict(zip(SIR_NAMES, p)) for p in pars]
    sys_ = lib.dust_system_create(lib.sir(), plist, 1, n_groups=len(plist), dt=dt, seed=seed)
    lib.dust_system_set_state_initial(sys_)
    out = lib.dust_system_simulate(sys_, time)
    return np.ascontiguousarray(out[:, :, 0, 4])


def contact_matrix(seed=7):
    rng = np.random.default_rng(seed)
    C = rng.uniform(0, 2, (16, 16))
    C = (C + C.T) / 2
    return C / np.max(np.abs(np.linalg.eigvals(C)))


def seir_groups(n_groups, seed=43):
    rng = np.random.default_rng(seed)
    C = contact_matrix()
    N = np.round(rng.uniform(1e4, 1e5, 16))
    return [{"beta": rng.uniform(0.1, 0.5), "sigma": 0.25, "gamma": rng.uniform(0.05, 0.2),
             "exp_noise": 1e6, "I0": 10.0, "N": N, "C": C} for _ in range(n_groups)]


def simulate_seir_data(lib, plist, time, dt, seed):
    """Weekly incidence of each age group from one stochastic run per group:
    a one-particle filter with no observations saves its trajectory."""
    nan = np.full((time.size, 16), np.nan)
    obj = lib.dust_filter_create(lib.seir16(), 0.0, {"time": time, "incidence": nan}, 1,
                                 n_groups=len(plist), dt=dt, seed=seed, precision="fp32")
    lib.dust_likelihood_run(obj, plist, save_trajectories=True)
    tr = lib.dust_likelihood_last_trajectories(obj)          # (T, G, 1, 80)
    return np.ascontiguousarray(tr[:, :, 0, 64:80])            # (T, G, 16)


def traffic_entry(key, kernel):
    """The committed counters of this workload's launch shape for `kernel`
    (profiles/hbm_traffic.json: one entry, or a list of entries when the
    autotuner may choose between kernels), or None."""
    try:
        with open(os.path.join(ROOT, "profiles", "hbm_traffic.json")) as f:
            tab = json.load(f)
    except OSError:
        return None, None
    entries = tab.get(key)
    if entries is None:
        return tab, None
    if isinstance(entries, dict):
        entries = [entries]
    for e in entries:
        if e["kernel"].startswith(kernel):
            return tab, e
    return tab, None


def measured_traffic(key, kernel):
    """HBM bytes per launch from the committed FETCH_SIZE / WRITE_SIZE
    profile of this workload and launch shape, or None."""
    tab, e = traffic_entry(key, kernel)
    if e is None:
        return None
    return 1024 * (tab["fetch_scale"] * e["fetch_kib"] + tab["write_scale"] * e["write_kib"])


_VALU_PEAK = {}


def valu_peak(fp64):
    """Wave-instructions per SIMD per ns of independent fused multiply-adds
    at 8 wavefronts per SIMD, measured now (dust2_amd_valu_probe), and the
    device's SIMD count."""
    if fp64 not in _VALU_PEAK:
        import ctypes as C
        from dust2_amd._lib import call
        ms, rate = np.zeros(1), np.zeros(1)
        n_simd = np.zeros(1, dtype=np.int32)
        call("dust2_amd_valu_probe", 0 if fp64 else 1, 8, 20000,
             ms.ctypes.data_as(C.POINTER(C.c_double)),
             rate.ctypes.data_as(C.POINTER(C.c_double)),
             n_simd.ctypes.data_as(C.POINTER(C.c_int)))
        _VALU_PEAK[fp64] = (float(rate[0]), int(n_simd[0]))
    return _VALU_PEAK[fp64]


def valu_issue(key, kernel, k_ms, fp64):
    """The dominant kernel's vector-instruction issue rate: its VALU
    instructions per launch (the committed SQ_INSTS_VALU profile of this
    workload and launch shape) over its event-timed duration, against the
    measured issue rate of independent FMAs; None without a profile."""
    _, e = traffic_entry(key, kernel)
    if e is None or "valu_instructions" not in e:
        return None
    peak, n_simd = valu_peak(fp64)
    achieved = e["valu_instructions"] / n_simd / (k_ms * 1e6)
    return {"instructions_per_launch": e["valu_instructions"], "achieved": achieved,
            "peak": peak, "unit": "wave-instructions/SIMD/ns", "frac": achieved / peak,
            "peak_of": "FP64 FMA" if fp64 else "FP32 FMA", "simds": n_simd}


def roofline(achieved, kernel, k_ms, b_launch, b_obs, traffic_key=None, fp64=False):
    return {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": measured_traffic(traffic_key, kernel) if traffic_key else None,
            "kernel": kernel, "kernel_ms": k_ms, "bytes_per_launch": b_launch,
            "bytes_per_particle_obs": b_obs,
            "valu_issue": valu_issue(traffic_key, kernel, k_ms, fp64) if traffic_key else None}


def timed_runs(lib, obj, plist, steps, warmup):
    for _ in range(warmup):
        lib.dust_likelihood_run(obj, plist)
    gpu_ms = []
    t0 = _time.perf_counter()
    for _ in range(steps):
        ll = lib.dust_likelihood_run(obj, plist)
        gpu_ms.append(lib.dust_likelihood_last_stats(obj)["kernel_ms"])
    return np.atleast_1d(ll), _time.perf_counter() - t0, gpu_ms


def pipeline_shape(obj, engine):
    """Launch chains and resample workgroup width of a pipeline-engine filter
    (the library's defaults, 2 and 1024, unless set or autotuned)."""
    if engine != "pipeline":
        return None
    chains = getattr(obj, "chains", None)
    block = getattr(obj, "resample_block", None)
    return {"chains": chains or 2, "resample_block": block or 1024}


def filter_roofline(lib, obj, plist, G, P, n_obs, S, s_x, engine_used, run_ms, key=None):
    b_obs = 2 * S * s_x + 2 * 32 + 2 * 8 + 2 * 4
    real = "float" if s_x == 4 else "double"
    if engine_used == "persistent":
        b_launch = G * P * n_obs * b_obs
        kernel = "filter_kernel<sir_model, %s, %d, %d>" % (
            (real,) + lib.dust_likelihood_last_stats(obj)["launch"])
        return roofline(b_launch / (run_ms * 1e-3) / 1e9, kernel, run_ms, b_launch, b_obs,
                        key, fp64=s_x == 8)
    # pipeline: the advance kernel moves state, RNG, weight write and
    # ancestor read of one interval; timed per launch in one profiled run
    obj.set_engine("pipeline", graph=False, profile=True)
    lib.dust_likelihood_run(obj, plist)
    k_ms = lib.dust_likelihood_last_stats(obj)["advance_ms"]
    obj.set_engine("pipeline")
    b_launch = G * P * (2 * S * s_x + 2 * 32 + 8 + 4)
    # the kernel's template arguments: model, real, workgroup size and the
    # minimum waves per SIMD of its launch bounds (discrete_kernels.hip)
    if S > 5:
        name = "advance_kernel<seir16_model, %s, %s>" % (real, "64, 3" if s_x == 4 else "256, 1")
    else:
        name = "advance_kernel<sir_model, %s, 256, 1>" % real
    out = roofline(b_launch / (k_ms * 1e-3) / 1e9, name, k_ms, b_launch, b_obs, key,
                   fp64=s_x == 8)
    # the whole run (advance + resample launches) against the same bytes
    out["run_ms"] = run_ms
    out["run_achieved"] = G * P * n_obs * b_obs / (run_ms * 1e-3) / 1e9
    return out


# ---------------------------------------------------------------------------
# workloads
# ---------------------------------------------------------------------------
def bench_sir(lib, args):
    dt = 0.25
    G, P = args.groups, args.particles
    pars = sir_groups(G)
    time = np.arange(1.0, args.n_obs + 1)
    data = simulate_sir_data(lib, pars, time, dt, seed=42)
    obj = lib.dust_filter_create(lib.sir(), 0.0, {"time": time, "incidence": data}, P,
                                 n_groups=G, dt=dt, seed=1, precision=args.precision)
    plist = [dict(zip(SIR_NAMES, p)) for p in pars]
    tuned = None
    if args.launch:
        obj.set_launch(*[int(x) for x in args.launch.split(",")])
    obj.set_engine(args.engine, graph=not args.no_graph)
    if args.engine == "auto" and not args.launch and not args.no_tune:
        # untimed: the autotuner runs each engine / launch shape three times and
        # restores the RNG streams, then keeps the fastest
        tuned = obj.autotune(plist)
    # every call passes the parameters, as an MCMC driver does
    ll, elapsed, gpu_ms = timed_runs(lib, obj, plist, args.steps, args.warmup)
    n_steps = int(round(time[-1] / dt))
    value = G * P * n_steps * args.steps / elapsed
    stats = lib.dust_likelihood_last_stats(obj)
    s_x = 4 if args.precision == "fp32" else 8
    rl = filter_roofline(lib, obj, plist, G, P, time.size, 5, s_x, stats["engine"],
                         float(np.mean(gpu_ms)), key=f"sir:{G}x{P}:{args.precision}")
    out = {
        "metric": "particle-steps/s", "value": value, "unit": "particle-steps/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32" if args.precision == "fp32" else "f64",
        "data": "synthetic (seeded simulation of each group's model)",
        "config": {"workload": "batched stochastic SIR bootstrap filter",
                   "groups": G, "particles": P, "n_obs": int(time.size), "dt": dt,
                   "precision": args.precision, "engine": stats["engine"],
                   "launch": stats["launch"] if stats["engine"] == "persistent" else None,
                   "pipeline": pipeline_shape(obj, stats["engine"])},
        "autotune": tuned,
        "likelihood_evals_per_s": G * args.steps / elapsed,
        "gpu_ms_per_step": float(np.mean(gpu_ms)),
        "mean_ll": float(np.mean(ll)),
        "roofline": rl,
    }
    # the same workload in FP64 (the reference's precision)
    if not args.no_fp64:
        obj64 = lib.dust_filter_create(lib.sir(), 0.0, {"time": time, "incidence": data}, P,
                                       n_groups=G, dt=dt, seed=1, precision="fp64")
        if not args.no_tune:
            obj64.autotune(plist)  # untimed, streams restored
        _, el64, ms64 = timed_runs(lib, obj64, plist, 2, 1)
        out["fp64"] = {"particle_steps_per_s": G * P * n_steps * 2 / el64,
                       "ms_per_run": 1e3 * el64 / 2, "gpu_ms_per_run": float(np.mean(ms64)),
                       "engine": lib.dust_likelihood_last_stats(obj64)["engine"],
                       "pipeline": pipeline_shape(obj64,
                                                  lib.dust_likelihood_last_stats(obj64)["engine"])}
        del obj64
    # the plain path: one filter object per group, one group per launch
    # sequence (warmed by one run each, then one timed run each)
    n_plain = 8
    plain_s = 0.0
    for g in range(n_plain):
        one = lib.dust_filter_create(lib.sir(), 0.0, {"time": time, "incidence": data[:, g]},
                                     P, dt=dt, seed=100 + g, precision=args.precision)
        lib.dust_likelihood_run(one, plist[g])
        t0 = _time.perf_counter()
        lib.dust_likelihood_run(one, plist[g])
        plain_s += _time.perf_counter() - t0
    out["plain_evals_per_s"] = n_plain / plain_s
    out["batched_over_plain"] = out["likelihood_evals_per_s"] / out["plain_evals_per_s"]
    if not args.no_cpu:
        ol = oracle()

        def cpu_run(g, k, n_threads):
            return ol.sir_filter(pars[:g], P, 0.0, time, dt, data[:, :g], seed=1 + k,
                                 n_threads=n_threads)["ll"]

        # about 3 s per run at nproc threads, 2 s at one; median of 3 (seeds 1, 2, 3)
        out["cpu_baseline"], g, ref = cpu_protocol(
            cpu_run, G, P * n_steps, "particle-steps/s", 3,
            lambda n: f"{n} of {G} groups x {P} particles x {time.size} obs")
        # FP32 batched filter against the FP64 oracle on the sampled groups:
        # K GPU replicates of each (one batched launch of K * g groups)
        # against the oracle's 3 replicates; z = sum of the per-group mean
        # differences over its Monte Carlo standard error, with each group's
        # variance from its GPU replicates
        K = 16
        rep = lib.dust_filter_create(lib.sir(), 0.0,
                                     {"time": time, "incidence": np.repeat(data[:, :g], K, 1)},
                                     P, n_groups=g * K, dt=dt, seed=2, precision=args.precision)
        ll_rep = np.atleast_1d(lib.dust_likelihood_run(rep, [p for p in plist[:g]
                                                             for _ in range(K)]))
        del rep
        ll_rep = ll_rep.reshape(g, K)
        m_gpu = ll_rep.mean(axis=1)
        v = ll_rep.var(axis=1, ddof=1)
        m_cpu = np.mean(np.array(ref), axis=0)
        se_g = np.sqrt(v / K + v / len(ref))
        out["vs_fp64_oracle"] = {
            "groups": g, "gpu_replicates": K, "oracle_replicates": len(ref),
            "mean_delta_ll": float(np.mean(m_gpu - m_cpu)),
            "delta_ll_in_mc_se": float(np.sum(m_gpu - m_cpu) / np.sqrt(np.sum(se_g ** 2))),
            "max_abs_group_z": float(np.max(np.abs(m_gpu - m_cpu) / se_g))}
    return out


def bench_sir_replicates(lib, args):
    dt = 0.25
    G, P = 1000, 1000
    row = np.array([1000.0, 10.0, 0.2, 0.1, 1e6])
    time = np.arange(1.0, 101.0)
    data = simulate_sir_data(lib, row[None, :], time, dt, seed=42)[:, 0]
    prec = args.precision if args.precision_set else "fp64"
    obj = lib.dust_filter_create(lib.sir(), 0.0, {"time": time, "incidence": data}, P,
                                 n_groups=G, dt=dt, seed=1, precision=prec)
    plist = [dict(zip(SIR_NAMES, row))] * G
    if args.launch:
        obj.set_launch(*[int(x) for x in args.launch.split(",")])
    if args.engine != "auto" or args.launch:
        obj.set_engine(args.engine, graph=not args.no_graph)
    tuned = None if (args.no_tune or args.launch or args.engine != "auto") else \
        obj.autotune(plist)
    ll, elapsed, gpu_ms = timed_runs(lib, obj, plist, args.steps, args.warmup)
    n_steps = 400
    stats = lib.dust_likelihood_last_stats(obj)
    s_x = 8 if prec == "fp64" else 4
    rl = filter_roofline(lib, obj, plist, G, P, time.size, 5, s_x, stats["engine"],
                         float(np.mean(gpu_ms)), key=f"sir-replicates:{G}x{P}:{prec}")
    out = {"metric": "likelihood evals/s", "value": G * args.steps / elapsed,
           "unit": "evals/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": 1e3 * elapsed / args.steps, "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f64" if prec == "fp64" else "f32",
           "data": "synthetic (100 daily observations simulated at seed 42)",
           "config": {"workload": "1000 replicate SIR filters, 1000 particles", "groups": G,
                      "particles": P, "n_obs": 100, "dt": dt, "precision": prec},
           "engine": stats["engine"], "autotune": tuned,
           "particle_steps_per_s": G * P * n_steps * args.steps / elapsed,
           "ll_mean": float(ll.mean()), "ll_var": float(ll.var(ddof=1)),
           "ll_mc_se": float(ll.std(ddof=1) / np.sqrt(G)), "roofline": rl}
    if not args.no_cpu:
        ol = oracle()

        def cpu_run(g, k, n_threads):
            return ol.sir_filter(np.tile(row, (g, 1)), P, 0.0, time, dt, data[:, None],
                                 seed=100 + k, n_threads=n_threads)["ll"]

        # about 3 s per run at nproc threads, 2 s at one; median of 5 (seeds 100 to 104)
        out["cpu_baseline"], g, runs = cpu_protocol(
            cpu_run, G, 1, "evals/s", 5, lambda n: f"{n} replicate filters")
        # the timed runs' streams (seed 1) are independent of the oracle's
        ref = np.concatenate(runs)
        se = float(np.sqrt(ll.var(ddof=1) / G + ref.var(ddof=1) / ref.size))
        out["oracle_replicates"] = int(ref.size)
        out["oracle_ll_mean"] = float(ref.mean())
        out["delta_ll_in_mc_se"] = float((ll.mean() - ref.mean()) / se)
        if prec == "fp64":
            # same streams: a fresh filter seeded as the oracle's first run
            # reproduces its g replicates draw for draw
            same = lib.dust_filter_create(lib.sir(), 0.0, {"time": time, "incidence": data}, P,
                                          n_groups=G, dt=dt, seed=100, precision=prec)
            ll_same = np.asarray(lib.dust_likelihood_run(same, plist))[:g]
            del same
            out["same_streams_max_rel_ll"] = float(np.max(np.abs(ll_same - runs[0])
                                                          / np.abs(runs[0])))
    return out


def bench_sirode(lib, args):
    n = 65536
    rng = np.random.default_rng(42)
    pars = np.column_stack([np.full(n, 1000.0), np.full(n, 10.0), rng.uniform(0.1, 0.5, n),
                            rng.uniform(0.05, 0.2, n)])
    time = np.arange(1.0, 366.0)
    data = simulate_sir_data(lib, np.array([[1000.0, 10.0, 0.2, 0.1, 1e6]]), time, 0.25,
                             seed=42)[:, 0]
    plist = [dict(zip(["N", "I0", "beta", "gamma"], p)) for p in pars]
    res = {}
    for prec in ["fp64", "fp32"]:
        obj = lib.dust_unfilter_create(lib.sirode(), 0.0, {"time": time, "incidence": data},
                                       n_particles=1, n_groups=n, precision=prec)
        for _ in range(args.warmup):
            obj.run(plist)
        ms, el = [], 0.0
        for k in range(args.steps):
            # with no warmup the first timed call also sets the parameters
            first = plist if (k == 0 and args.warmup == 0) else None
            t0 = _time.perf_counter()
            ll = obj.run(first)
            el += _time.perf_counter() - t0
            ms.append(obj.last_stats()[1])  # kernel time, read outside the timed run
        stats, _ = obj.last_stats()
        res[prec] = {"evals_per_s": n * args.steps / el, "wall_ms": el / args.steps * 1e3,
                     "kernel_ms": float(np.mean(ms)),
                     "attempts_per_s": float(stats[:, 0].sum()) / (np.mean(ms) * 1e-3),
                     "mean_steps": float(stats[:, 0].mean()), "ll": ll}
    # likelihood and gradient (beta, gamma, I0) by forward sensitivities, FP64
    gobj = lib.dust_unfilter_create(lib.sirode(), 0.0, {"time": time, "incidence": data},
                                    n_particles=1, n_groups=n)
    gobj.run(plist, adjoint=True)
    t0 = _time.perf_counter()
    for _ in range(args.steps):
        gobj.run(None, adjoint=True)
    grad_el = _time.perf_counter() - t0
    grad = gobj.last_gradient()
    out = {"metric": "likelihood evals/s", "value": res["fp64"]["evals_per_s"],
           "unit": "evals/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": res["fp64"]["wall_ms"], "higher_is_better": True, "scaling": "weak",
           "gradient_evals_per_s": n * args.steps / grad_el,
           "gradient_kernel_ms": gobj.last_stats()[1],
           "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded SIR simulation)",
           "config": {"workload": "SIR ODE unfilter, Dormand-Prince 5(4)", "parameter_sets": n,
                      "n_obs": 365, "rtol": 1e-6, "atol": 1e-6},
           "fp64": {k: v for k, v in res["fp64"].items() if k != "ll"},
           "fp32_state": {k: v for k, v in res["fp32"].items() if k != "ll"},
           "fp32_vs_fp64_max_rel_ll": float(np.max(np.abs(res["fp32"]["ll"] - res["fp64"]["ll"])
                                                   / np.abs(res["fp64"]["ll"])))}
    if not args.no_cpu:
        ol = oracle()
        threads = cpu_threads()

        def cpu_run(g, k, n_threads):
            return ol.sirode_unfilter(pars[:g], 1, 0.0, time, data, n_threads=n_threads)

        # up to all n sets per run at nproc threads (about 2 s at one); median of 5
        out["cpu_baseline"], m, runs = cpu_protocol(
            cpu_run, n, 1, "evals/s", 5, lambda g: f"{g} of {n} parameter sets")
        ref = runs[0]
        out["fp64_vs_oracle_max_rel_ll"] = float(np.max(np.abs(res["fp64"]["ll"][:m] - ref["ll"])
                                                        / np.abs(ref["ll"])))
        k = 4096
        _, ref_grad = ol.sirode_gradient(pars[:k], 0.0, time, data, n_threads=threads)
        out["gradient_vs_oracle_max_rel"] = float(np.max(np.abs(grad[:k] - ref_grad)
                                                         / np.abs(ref_grad)))
    return out


def bench_seir16(lib, args):
    dt = 0.25
    G, P = 128, 8192
    n_obs = args.n_obs_seir
    time = np.arange(7.0, 7.0 * (n_obs + 1), 7.0)
    plist = seir_groups(G)
    data = simulate_seir_data(lib, plist, time, dt, seed=42)
    obj = lib.dust_filter_create(lib.seir16(), 0.0, {"time": time, "incidence": data}, P,
                                 n_groups=G, dt=dt, seed=1, precision="fp32")
    # untimed: one to four launch chains, three runs each, streams restored
    tuned = None if args.no_tune else obj.autotune(plist)
    ll, elapsed, gpu_ms = timed_runs(lib, obj, plist, args.steps, args.warmup)
    n_steps = int(round(time[-1] / dt))
    value = G * P * n_steps * args.steps / elapsed
    rl = filter_roofline(lib, obj, plist, G, P, time.size, 80, 4, "pipeline",
                         float(np.mean(gpu_ms)), key=f"seir16:{G}x{P}:fp32")
    # plain path: one filter object per group, one group per run; each
    # object is warmed by one run, then a second run is timed
    n_plain = 4
    plain_s = 0.0
    for g in range(n_plain):
        one = lib.dust_filter_create(lib.seir16(), 0.0,
                                     {"time": time, "incidence": data[:, g, :]}, P, dt=dt,
                                     seed=100 + g, precision="fp32")
        lib.dust_likelihood_run(one, plist[g])
        t0 = _time.perf_counter()
        lib.dust_likelihood_run(one, plist[g])
        plain_s += _time.perf_counter() - t0
    plain_evals = n_plain / plain_s
    batched_evals = G * args.steps / elapsed
    out = {"metric": "particle-steps/s", "value": value, "unit": "particle-steps/s",
           "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": 1e3 * elapsed / args.steps, "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f32",
           "data": "synthetic (seeded simulation of each group's model)",
           "config": {"workload": "age-structured SEIR16 bootstrap filter (MFMA FOI)",
                      "groups": G, "particles": P, "n_obs": n_obs, "dt": dt,
                      "pipeline": pipeline_shape(obj, "pipeline")},
           "likelihood_evals_per_s": batched_evals, "plain_evals_per_s": plain_evals,
           "batched_over_plain": batched_evals / plain_evals, "autotune": tuned,
           "mean_ll": float(np.mean(ll)), "roofline": rl}
    if not args.no_cpu:
        ol = oracle()
        rows = np.array([np.concatenate([[q["beta"], q["sigma"], q["gamma"], q["exp_noise"],
                                          q["I0"]], q["N"], np.ravel(q["C"])])
                         for q in plist[:1]])

        def cpu_run(p, k, n_threads):
            return ol.sir_filter(rows, p, 0.0, time, dt, data[:, :1, :], seed=1 + k,
                                 n_threads=n_threads, model="seir16")

        # one group over all the weekly observations with p of its particles
        # (so that the sample follows the whole epidemic); up to P particles
        # per run at nproc threads, about 2 s at one; median of 3
        out["cpu_baseline"], _, _ = cpu_protocol(
            cpu_run, P, int(round(time[-1] / dt)), "particle-steps/s", 3,
            lambda p: f"1 of {G} groups x {p} of {P} particles x {time.size} weekly obs")
    return out


def bench_serving(lib, args):
    from dust2_amd.serving import Server
    n_chains = 64
    rate = 2000.0
    duration = args.serve_seconds
    C = contact_matrix()
    gens = {"sir": lib.sir(), "seir16": lib.seir16(), "sirode": lib.sirode()}

    def make_chains(seed):
        rng = np.random.default_rng(seed)
        chains = []
        for c in range(n_chains):
            model = ["sir", "seir16", "sirode"][c % 3]
            T = int(rng.integers(50, 731))
            P = int(rng.choice([128, 512, 2048, 8192])) if model != "sirode" else 1
            if model == "seir16":
                time = np.arange(7.0, 7.0 * (T + 1), 7.0)
                data = rng.poisson(30, (T, 16)).astype(float)
            else:
                time = np.arange(1.0, T + 1)
                data = rng.poisson(30 if model == "sir" else 6, T).astype(float)
            chains.append({"model": model, "T": T, "P": P, "index": c,
                           "data": {"time": time, "incidence": data},
                           "theta": np.array([0.3, 0.1]), "ll": None, "busy": False,
                           "obj": None})
        return chains

    def pars(ch, theta):
        beta, gamma = float(abs(theta[0])), float(abs(theta[1]))
        if ch["model"] == "seir16":
            return {"beta": beta, "sigma": 0.25, "gamma": gamma, "exp_noise": 1e6, "I0": 10.0,
                    "N": np.full(16, 2e4), "C": C}
        p = {"N": 1000.0, "I0": 10.0, "beta": beta, "gamma": gamma}
        if ch["model"] == "sir":
            p["exp_noise"] = 1e6
        return p

    def accept(rng, ch, prop, ll):
        if ch["ll"] is None or np.log(rng.random()) < ll - ch["ll"]:
            ch["theta"], ch["ll"] = prop, ll

    def summary(lat, soj, done, el, n_arrived):
        # lat: submission -> result; soj: Poisson arrival -> result, which
        # grows without bound once arrivals outpace the service rate
        lat = np.array(lat) if lat else np.array([np.nan])
        soj = np.array(soj) if soj else np.array([np.nan])
        return {"evals_per_s": done / el, "completed": done, "arrived": n_arrived,
                "p50_ms": float(np.percentile(lat, 50) * 1e3),
                "p99_ms": float(np.percentile(lat, 99) * 1e3),
                "sojourn_p50_ms": float(np.percentile(soj, 50) * 1e3),
                "sojourn_p99_ms": float(np.percentile(soj, 99) * 1e3)}

    def run_batched():
        # each arrival of the Poisson stream is one proposal of an idle chain
        rng = np.random.default_rng(11)
        chains = make_chains(5)
        arrivals = np.cumsum(rng.exponential(1 / rate, int(rate * duration * 2) + 10))
        sv = Server(hbm_bytes=64 << 30, chunk_intervals=args.chunk)
        ai, done, lat, soj, pending = 0, 0, [], [], {}
        t0 = _time.perf_counter()
        while True:
            now = _time.perf_counter() - t0
            if now > duration:
                break
            while ai < arrivals.size and arrivals[ai] <= now:
                idle = [c for c in chains if not c["busy"]]
                if not idle:
                    break
                ch = idle[int(rng.integers(len(idle)))]
                prop = ch["theta"] + rng.normal(0, 0.02, 2)
                # each chain's streams stay resident and continue from its
                # previous request, seeded at long_jump^chain (R/monty.R:29-36)
                rid = sv.submit(gens[ch["model"]], pars(ch, prop), ch["P"], 0.0, ch["data"],
                                dt=0.25, seed=11, chain=ch["index"])
                ch["busy"] = True
                pending[rid] = (ch, prop, arrivals[ai], _time.perf_counter() - t0)
                ai += 1
            if not pending:
                _time.sleep(1e-4)
                continue
            sv.step()
            t_now = _time.perf_counter() - t0
            for rid, ll, status, _ in sv.poll():
                ch, prop, t_arr, t_sub = pending.pop(rid)
                lat.append(t_now - t_sub)
                soj.append(t_now - t_arr)
                accept(rng, ch, prop, ll)
                ch["busy"] = False
                done += 1
        return summary(lat, soj, done, _time.perf_counter() - t0, ai)

    def run_plain():
        # the plain path: one filter (or unfilter) object per chain, each
        # request a synchronous dust_likelihood_run, one at a time
        rng = np.random.default_rng(11)
        chains = make_chains(5)
        arrivals = np.cumsum(rng.exponential(1 / rate, int(rate * duration * 2) + 10))
        ai, done, lat, soj = 0, 0, [], []
        t0 = _time.perf_counter()
        while True:
            now = _time.perf_counter() - t0
            if now > duration:
                break
            if arrivals[ai] > now:
                _time.sleep(min(arrivals[ai] - now, 1e-3))
                continue
            ch = chains[int(rng.integers(n_chains))]
            prop = ch["theta"] + rng.normal(0, 0.02, 2)
            if ch["obj"] is None:
                g = gens[ch["model"]]
                if ch["model"] == "sirode":
                    ch["obj"] = lib.dust_unfilter_create(g, 0.0, ch["data"])
                else:
                    ch["obj"] = lib.dust_filter_create(g, 0.0, ch["data"], ch["P"], dt=0.25,
                                                       seed=int(rng.integers(0, 2**62)),
                                                       precision="fp32")
            t_sub = _time.perf_counter()
            ll = lib.dust_likelihood_run(ch["obj"], pars(ch, prop))
            lat.append(_time.perf_counter() - t_sub)
            soj.append(_time.perf_counter() - t0 - arrivals[ai])
            accept(rng, ch, prop, ll)
            ai += 1
            done += 1
        return summary(lat, soj, done, _time.perf_counter() - t0,
                       int(np.sum(arrivals <= duration)))

    batched = run_batched()
    plain = run_plain()
    return {"metric": "likelihood evals/s", "value": batched["evals_per_s"], "unit": "evals/s",
            "n_gpus": 1, "steps": 1, "warmup": 0, "ms_per_step": duration * 1e3,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
            "data": "synthetic (random-walk MCMC proposals on seeded Poisson data)",
            "config": {"workload": "continuous-batching likelihood server", "chains": n_chains,
                       "arrival_rate_per_s": rate, "chunk_intervals": args.chunk,
                       "seconds": duration},
            "batched": batched, "plain_one_at_a_time": plain,
            "batched_over_plain": batched["evals_per_s"] / max(plain["evals_per_s"], 1e-9)}


def bench_valu_peak(lib, args):
    """The vector-ALU issue rate of independent fused multiply-adds at 1, 2,
    4 and 8 wavefronts per SIMD (dust2_amd_valu_probe), FP32 and FP64: the
    peak against which the sampler kernels' VALU instruction counts are
    set."""
    import ctypes as C
    from dust2_amd._lib import call
    out = {}
    for prec, code in (("fp32", 1), ("fp64", 0)):
        rows = []
        for w in (1, 2, 4, 8):
            ms = np.zeros(1)
            rate = np.zeros(1)
            n_simd = np.zeros(1, dtype=np.int32)
            call("dust2_amd_valu_probe", code, w, 20000,
                 ms.ctypes.data_as(C.POINTER(C.c_double)),
                 rate.ctypes.data_as(C.POINTER(C.c_double)),
                 n_simd.ctypes.data_as(C.POINTER(C.c_int)))
            rows.append({"waves_per_simd": w, "ms": float(ms[0]),
                         "wave_instr_per_simd_per_ns": float(rate[0])})
        out[prec] = rows
    peak = max(r["wave_instr_per_simd_per_ns"] for r in out["fp32"])
    return {"metric": "VALU wave-instructions per SIMD per ns", "value": peak,
            "unit": "wave-instr/SIMD/ns", "n_gpus": 1, "steps": 1, "warmup": 1,
            "ms_per_step": out["fp32"][-1]["ms"], "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f32", "data": "synthetic (independent FMA chains)",
            "config": {"workload": "vector-ALU issue-rate calibration"}, "rates": out}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="sir",
                    choices=["sir", "sir-replicates", "sirode", "seir16", "serving",
                             "valu-peak"])
    ap.add_argument("--groups", type=int, default=512)
    ap.add_argument("--particles", type=int, default=2048)
    ap.add_argument("--n-obs", type=int, default=365)
    ap.add_argument("--n-obs-seir", type=int, default=200)
    ap.add_argument("--precision", default=None)
    ap.add_argument("--launch", default=None, help="block,particles_per_thread")
    ap.add_argument("--engine", default="auto", choices=["auto", "persistent", "pipeline"])
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="skip the engine autotuner")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 run of the default "
                                                            "workload")
    ap.add_argument("--chunk", type=int, default=16)
    ap.add_argument("--serve-seconds", type=float, default=15.0)
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                                   "(parameter groups shard across GPUs with no "
                                   "communication during a run)"}))
        sys.exit(2)
    args.precision_set = args.precision is not None
    if args.precision is None:
        args.precision = "fp32"

    import dust2_amd as lib
    from dust2_amd._lib import call
    call("dust2_amd_init")
    fn = {"sir": bench_sir, "sir-replicates": bench_sir_replicates, "sirode": bench_sirode,
          "seir16": bench_seir16, "serving": bench_serving,
          "valu-peak": bench_valu_peak}[args.workload]
    print(json.dumps(fn(lib, args)))


if __name__ == "__main__":
    main()
