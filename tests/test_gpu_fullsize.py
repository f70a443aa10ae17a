"""Parity at the benchmark workloads' full sizes, through properties that do
not depend on size.

* Batched SIR filter, 512 groups x 2048 particles x 365 daily observations
  (dt = 0.25): the FP64 run is compared draw-for-draw with the oracle on
  sampled groups (each group is an independent filter over its own block of
  RNG streams, exported from the device before the run); the FP32 run is
  compared with the FP64 run group by group (paired differences centred on
  zero), over 32 replicates of every group in each precision, and, for the
  largest- and smallest-population groups, over 1000 replicates in each
  precision; every final state conserves the population exactly.
* SIR ODE unfilter, 65,536 parameter sets x 365 observations: every
  log-likelihood and final state against the oracle.
* Age-structured SEIR16 filter, 128 groups x 8192 particles x 200 weekly
  observations: population conserved exactly in every age group of every
  particle, incidence non-negative, finite log-likelihoods.
"""
import numpy as np
import pytest

import oracle_lib as ol

pytestmark = pytest.mark.gpu

SIR_NAMES = ["N", "I0", "beta", "gamma", "exp_noise"]


def sir_groups(G, seed=42):
    rng = np.random.default_rng(seed)
    beta = rng.uniform(0.1, 0.5, G)
    gamma = rng.uniform(0.05, 0.2, G)
    N = np.round(rng.uniform(1e3, 1e6, G))
    return np.column_stack([N, np.full(G, 10.0), beta, gamma, np.full(G, 1e6)])


def sir_data(pars, time, seed):
    out, _ = ol.sir_simulate(pars, 1, 0.0, 0.25, time, ol.prng(len(pars), seed=seed),
                             n_threads=8)
    return np.ascontiguousarray(out[:, :, 0, 4])             # (n_time, G)


@pytest.fixture(scope="module")
def sir_c2(gpu_lib):
    G, P = 512, 2048
    time = np.arange(1.0, 366.0)
    pars = sir_groups(G)
    data = sir_data(pars, time, 7)
    plist = [dict(zip(SIR_NAMES, p)) for p in pars]
    out = {"pars": pars, "data": data, "time": time, "P": P}
    for prec in ["fp64", "fp32"]:
        obj = gpu_lib.dust_filter_create(gpu_lib.sir(), 0.0, {"time": time, "incidence": data},
                                         P, n_groups=G, dt=0.25, seed=11, precision=prec)
        rng0 = gpu_lib.dust_likelihood_rng_state(obj).reshape(-1, 4).copy()
        ll = np.atleast_1d(gpu_lib.dust_likelihood_run(obj, plist))
        out[prec] = {"ll": ll, "rng0": rng0, "state": gpu_lib.dust_likelihood_last_state(obj),
                     "engine": gpu_lib.dust_likelihood_last_stats(obj)["engine"]}
    return out


@pytest.mark.parametrize("g", [0, 137, 511])
def test_sir_full_size_fp64_groups_match_oracle(sir_c2, g):
    P = sir_c2["P"]
    r = sir_c2["fp64"]
    block = r["rng0"][g * (P + 2):(g + 1) * (P + 2)]
    ref = ol.sir_filter(sir_c2["pars"][g:g + 1], P, 0.0, sir_c2["time"], 0.25,
                        sir_c2["data"][:, g:g + 1], rng=block, n_threads=8)
    np.testing.assert_allclose(r["ll"][g], ref["ll"][0], rtol=1e-12)
    assert np.array_equal(r["state"][g], ref["state"][0])


def test_sir_full_size_population_conserved(sir_c2):
    N = sir_c2["pars"][:, 0]
    for prec in ["fp64", "fp32"]:
        st = sir_c2[prec]["state"]                            # (G, P, 5)
        assert np.all(st[:, :, :3] >= 0)
        assert np.array_equal(st[:, :, 0] + st[:, :, 1] + st[:, :, 2],
                              np.broadcast_to(N[:, None], st.shape[:2]))
        assert np.all(st[:, :, 3] >= st[:, :, 4])
        assert np.all(np.isfinite(sir_c2[prec]["ll"]))


def test_sir_full_size_fp32_paired_with_fp64(sir_c2):
    # the two precisions draw different random numbers; over 512 groups the
    # paired differences of the log-likelihood estimates centre on zero
    # (bar 2 standard errors, p = 0.05; the seeds are fixed and the measured
    # z is -1.66; the replicate test below resolves each group's bias)
    d = sir_c2["fp32"]["ll"] - sir_c2["fp64"]["ll"]
    se = d.std(ddof=1) / np.sqrt(d.size)
    assert abs(d.mean()) < 2 * se, (d.mean(), se)


def test_sir_full_size_fp32_unbiased_all_groups(gpu_lib):
    # every one of the 512 groups, 32 FP32 replicate filters against 32 of
    # the FP64 filter (the oracle draw for draw), each replicate on its own
    # streams: the summed mean differences over their standard error below
    # 3.3 (p = 0.001), the per-group z spread near 1 (a bias or a variance
    # change confined to some parameter regimes widens it), and the mean
    # log ratio of the two variances within 4 standard errors of zero.
    # Measured: summed z 1.33, per-group z standard deviation 1.03 (5.7 %
    # beyond 2), log variance ratio -0.017 +- 0.018.
    G, K, P = 512, 32, 2048
    time = np.arange(1.0, 366.0)
    pars = sir_groups(G)
    data = np.repeat(sir_data(pars, time, 7), K, axis=1)
    plist = [dict(zip(SIR_NAMES, p)) for p in pars for _ in range(K)]
    ll = {}
    for prec, seed in (("fp32", 303), ("fp64", 404)):
        obj = gpu_lib.dust_filter_create(gpu_lib.sir(), 0.0, {"time": time, "incidence": data},
                                         P, n_groups=G * K, dt=0.25, seed=seed, precision=prec)
        ll[prec] = np.atleast_1d(gpu_lib.dust_likelihood_run(obj, plist)).reshape(G, K)
        del obj
    a, b = ll["fp32"], ll["fp64"]
    va, vb = a.var(axis=1, ddof=1), b.var(axis=1, ddof=1)
    se = np.sqrt(va / K + vb / K)
    d = a.mean(axis=1) - b.mean(axis=1)
    z_sum = d.sum() / np.sqrt(np.sum(se ** 2))
    z = d / se
    lr = np.log(va / vb)
    print(f"z_sum {z_sum:.3f} z_sd {z.std():.3f} frac|z|>2 {np.mean(np.abs(z) > 2):.4f} "
          f"log var ratio {lr.mean():.4f} +- {lr.std(ddof=1) / np.sqrt(G):.4f}")
    assert abs(z_sum) < 3.3, z_sum
    assert 0.88 < z.std() < 1.15, z.std()
    assert abs(lr.mean()) < 4 * lr.std(ddof=1) / np.sqrt(G), lr.mean()


def test_sir_full_size_fp32_unbiased_per_group(gpu_lib):
    # 1000 replicate filters (2048 particles, 365 days) of the benchmark's
    # two largest-population groups (N ~ 1e6, where FP32 resolves the state
    # least finely) and its two smallest, in FP32 and in FP64 (the FP64 GPU
    # filter is the oracle draw for draw), each replicate on its own
    # streams.  Bars: per group |z| of the mean difference below 3 and the
    # sum of z^2 below the 99.9 % point of chi^2(4); variance ratios inside
    # the two-sided 99.9 % F(999, 999) interval.  Measured over twelve
    # groups of this workload: |z| <= 2.03, no consistent sign, variance
    # ratios 0.93-1.12.
    from scipy import stats
    G, REPS, P = 512, 1000, 2048
    time = np.arange(1.0, 366.0)
    pars_all = sir_groups(G)
    order = np.argsort(pars_all[:, 0])
    pick = [order[-1], order[-2], order[0], order[1]]
    data_all = sir_data(pars_all[pick], time, 7)
    z = []
    lo, hi = stats.f.ppf([0.0005, 0.9995], REPS - 1, REPS - 1)
    for i, g in enumerate(pick):
        pars = [dict(zip(SIR_NAMES, pars_all[g]))] * REPS
        data = np.repeat(data_all[:, i:i + 1], REPS, axis=1)
        ll = {}
        for prec, seed in (("fp32", 101), ("fp64", 202)):
            obj = gpu_lib.dust_filter_create(gpu_lib.sir(), 0.0, {"time": time, "incidence": data},
                                             P, n_groups=REPS, dt=0.25, seed=seed,
                                             precision=prec)
            ll[prec] = np.atleast_1d(gpu_lib.dust_likelihood_run(obj, pars))
            del obj
        a, b = ll["fp32"], ll["fp64"]
        se = np.sqrt(a.var(ddof=1) / REPS + b.var(ddof=1) / REPS)
        z.append((a.mean() - b.mean()) / se)
        assert abs(z[-1]) < 3, (g, a.mean(), b.mean(), se)
        assert lo < a.var(ddof=1) / b.var(ddof=1) < hi, g
    assert np.sum(np.square(z)) < stats.chi2.ppf(0.999, len(z)), z


def test_sirode_full_size_matches_oracle(gpu_lib):
    rng = np.random.default_rng(5)
    n = 65536
    pars = np.column_stack([np.full(n, 1000.0), np.full(n, 10.0), rng.uniform(0.1, 0.5, n),
                            rng.uniform(0.05, 0.2, n)])
    time = np.arange(1.0, 366.0)
    data = rng.poisson(6, 365).astype(float)
    obj = gpu_lib.dust_unfilter_create(gpu_lib.sirode(), 0.0, {"time": time, "incidence": data},
                                       n_particles=1, n_groups=n)
    ll = np.atleast_1d(obj.run([dict(zip(["N", "I0", "beta", "gamma"], p)) for p in pars]))
    ref = ol.sirode_unfilter(pars, 1, 0.0, time, data, n_threads=8)
    np.testing.assert_allclose(ll, ref["ll"], rtol=1e-10)
    np.testing.assert_allclose(obj.last_state(), ref["state"], rtol=1e-10, atol=1e-9)


def contact_matrix(seed=7):
    rng = np.random.default_rng(seed)
    C = rng.uniform(0, 2, (16, 16))
    C = (C + C.T) / 2
    return C / np.max(np.abs(np.linalg.eigvals(C)))


def test_seir16_full_size_invariants(gpu_lib):
    G, P = 128, 8192
    rng = np.random.default_rng(43)
    C = contact_matrix()
    N = np.round(rng.uniform(1e4, 1e5, 16))
    plist = [{"beta": rng.uniform(0.1, 0.5), "sigma": 0.25, "gamma": rng.uniform(0.05, 0.2),
              "exp_noise": 1e6, "I0": 10.0, "N": N, "C": C} for _ in range(G)]
    time = 7.0 * np.arange(1, 201)
    # data: one stochastic run of each group (a one-particle filter with no
    # observations saves its trajectory)
    nan = np.full((time.size, 16), np.nan)
    sim = gpu_lib.dust_filter_create(gpu_lib.seir16(), 0.0, {"time": time, "incidence": nan}, 1,
                                     n_groups=G, dt=0.25, seed=1, precision="fp32")
    gpu_lib.dust_likelihood_run(sim, plist, save_trajectories=True)
    data = np.ascontiguousarray(gpu_lib.dust_likelihood_last_trajectories(sim)[:, :, 0, 64:80])
    obj = gpu_lib.dust_filter_create(gpu_lib.seir16(), 0.0, {"time": time, "incidence": data},
                                     P, n_groups=G, dt=0.25, seed=2, precision="fp32")
    ll = np.atleast_1d(gpu_lib.dust_likelihood_run(obj, plist))
    assert ll.shape == (G,) and np.all(np.isfinite(ll))
    st = gpu_lib.dust_likelihood_last_state(obj)             # (G, P, 80)
    pop = st[:, :, 0:16] + st[:, :, 16:32] + st[:, :, 32:48] + st[:, :, 48:64]
    assert np.array_equal(pop, np.broadcast_to(N, pop.shape))
    assert np.all(st[:, :, :64] >= 0) and np.all(st[:, :, 64:80] >= 0)
