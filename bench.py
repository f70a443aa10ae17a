"""Benchmark: batched bootstrap particle filter on one MI355X.

Workload (default): 512 parameter groups x 2048 particles of the stochastic
SIR model (beta ~ U(0.1, 0.5), gamma ~ U(0.05, 0.2), N ~ U(1e3, 1e6)
integer, I0 = 10), dt = 0.25, 365 daily Poisson observations per group
simulated from the model with a fixed seed, FP32 state and samplers with
FP64 log-weights, resampling sums and likelihood.  One step = one complete
filter run of every group (1.53e9 particle-steps).

Prints one JSON line: particle-steps/s, the roofline figure of the filter
kernel (algorithmic bytes per particle per observation, over the kernel's
own event-timed duration) and the CPU oracle timed on a bounded sample.
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


def make_workload(n_groups, n_time, seed=42):
    rng = np.random.default_rng(seed)
    beta = rng.uniform(0.1, 0.5, n_groups)
    gamma = rng.uniform(0.05, 0.2, n_groups)
    N = np.round(rng.uniform(1e3, 1e6, n_groups))
    pars = np.column_stack([N, np.full(n_groups, 10.0), beta, gamma,
                            np.full(n_groups, 1e6)])
    return pars, np.arange(1.0, n_time + 1)


def simulate_data(lib, pars, time, dt, seed):
    """Daily incidence from one stochastic run of each group's own model
    (GPU system kernels)."""
    plist = [dict(zip(SIR_NAMES, p)) for p in pars]
    sys_ = lib.dust_system_create(lib.sir(), plist, 1, n_groups=len(plist), dt=dt, seed=seed)
    lib.dust_system_set_state_initial(sys_)
    out = lib.dust_system_simulate(sys_, time)
    return np.ascontiguousarray(out[:, :, 0, 4])


def cpu_baseline(pars, time, dt, data, n_particles, target_s=15.0):
    """FP64 CPU oracle (OpenMP over group x particle, as dust2) on a bounded
    sample of the same workload: as many groups as fit ~target_s."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as ol
    threads = min(16, os.cpu_count() or 1)
    g = 1
    while True:
        t0 = _time.perf_counter()
        ol.sir_filter(pars[:g], n_particles, 0.0, time, dt, data[:, :g], seed=1,
                      n_threads=threads)
        el = _time.perf_counter() - t0
        if el > target_s / 4 or g >= len(pars):
            break
        g = min(len(pars), max(g * 2, int(g * target_s / 4 / max(el, 1e-3))))
    steps = g * n_particles * round((time[-1] - 0.0) / dt)
    return {"value": steps / el, "unit": "particle-steps/s", "cores": threads,
            "kind": "port",
            "sample": f"{g} of {len(pars)} groups x {n_particles} particles x {time.size} obs, "
                      f"FP64 oracle, {el:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--groups", type=int, default=512)
    ap.add_argument("--particles", type=int, default=2048)
    ap.add_argument("--n-obs", type=int, default=365)
    ap.add_argument("--precision", default="fp32")
    ap.add_argument("--launch", default=None, help="block,particles_per_thread")
    ap.add_argument("--engine", default="auto", choices=["auto", "persistent", "pipeline"])
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                                   "(groups shard across GPUs with no communication, "
                                   "see DESIGN.md)"}))
        sys.exit(2)

    import dust2_amd as lib
    from dust2_amd._lib import call
    call("dust2_amd_init")

    dt = 0.25
    G, P = args.groups, args.particles
    pars, time = make_workload(G, args.n_obs)
    data = simulate_data(lib, pars, time, dt, seed=42)
    obj = lib.dust_filter_create(lib.sir(), 0.0, {"time": time, "incidence": data}, P,
                                 n_groups=G, dt=dt, seed=1, precision=args.precision)
    if args.launch:
        obj.set_launch(*[int(x) for x in args.launch.split(",")])
    obj.set_engine(args.engine, graph=not args.no_graph)
    plist = [dict(zip(SIR_NAMES, p)) for p in pars]
    # every call passes the parameters, as an MCMC driver does (the update
    # is host-side bookkeeping)
    for _ in range(args.warmup):
        lib.dust_likelihood_run(obj, plist)
    kernel_ms = []
    t0 = _time.perf_counter()
    for _ in range(args.steps):
        ll = lib.dust_likelihood_run(obj, plist)
        kernel_ms.append(lib.dust_likelihood_last_stats(obj)["kernel_ms"])
    elapsed = _time.perf_counter() - t0
    ms_per_step = 1e3 * elapsed / args.steps
    n_steps = int(round(time[-1] / dt))
    psteps = G * P * n_steps
    value = psteps * args.steps / elapsed
    s_x = 4 if args.precision == "fp32" else 8
    # algorithmic bytes per particle per observation: state gather-read and
    # write, xoshiro256+ read and write, FP64 weight (lambda) write and read,
    # u32 ancestor write and read
    b_obs = 2 * 5 * s_x + 2 * 32 + 2 * 8 + 2 * 4
    stats = lib.dust_likelihood_last_stats(obj)
    engine = stats["engine"]
    if engine == "persistent":
        kernel = "sir_filter_kernel"
        k_ms = float(np.mean(kernel_ms))
        b_launch = G * P * time.size * b_obs
    else:
        # the advance kernel moves the state, RNG, lambda write and
        # ancestor read of one interval; timed per launch with events in
        # one extra profiled run
        kernel = "sir_advance_kernel"
        obj.set_engine("pipeline", graph=False, profile=True)
        lib.dust_likelihood_run(obj, plist)
        k_ms = lib.dust_likelihood_last_stats(obj)["advance_ms"]
        obj.set_engine(args.engine, graph=not args.no_graph)
        b_launch = G * P * (2 * 5 * s_x + 2 * 32 + 8 + 4)
    achieved = b_launch / (k_ms * 1e-3) / 1e9
    out = {
        "metric": "particle-steps/s", "value": value, "unit": "particle-steps/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32" if args.precision == "fp32" else "f64",
        "data": "synthetic (seeded simulation of each group's model)",
        "config": {"workload": "batched stochastic SIR bootstrap filter",
                   "groups": G, "particles": P, "n_obs": int(time.size), "dt": dt,
                   "precision": args.precision, "launch": obj.launch, "engine": engine,
                   "graph": not args.no_graph},
        "likelihood_evals_per_s": G * args.steps / elapsed,
        "gpu_ms_per_step": float(np.mean(kernel_ms)),
        "mean_ll": float(np.mean(ll)),
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": kernel, "kernel_ms": k_ms, "bytes_per_launch": b_launch,
                     "bytes_per_particle_obs": b_obs},
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(pars, time, dt, data, P)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
