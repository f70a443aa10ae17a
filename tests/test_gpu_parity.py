"""GPU parity: the HIP kernels, called through the C ABI, against the CPU
oracle on the same seeded inputs.

Bars: RNG streams and resampling indices bit-exact; FP64 filter states,
ancestors and RNG state identical draw-for-draw, log-likelihood within
1e-12 relative (the only differences allowed are last-place differences of
device vs host exp/log); FP32 filter statistically equal to the FP64 oracle
(over 1000 replicates each: mean log-likelihood within 1 Monte Carlo
standard error, variances and distributions indistinguishable).
"""
import numpy as np
import pytest

import oracle_lib as ol

pytestmark = pytest.mark.gpu

SIR = [1000.0, 10.0, 0.2, 0.1, 1e6]


def sir_data(n_time, seed=42, pars=SIR, dt=0.25):
    """Daily incidence simulated from the model itself (one particle)."""
    rng = ol.prng(1, seed=seed)
    out, _ = ol.sir_simulate(np.array([pars]), 1, 0.0, dt, np.arange(1.0, n_time + 1), rng)
    return out[:, 0, 0, 4].copy()


# ---------------------------------------------------------------------------
def test_rng_streams_bit_exact(gpu_lib):
    for seed in [0, 42, 2**40 + 3]:
        got = gpu_lib.gpu_rng_streams(3000, seed=seed)
        assert np.array_equal(got, ol.prng(3000, seed=seed))
    raw = ol.prng(3, seed=9)
    got = gpu_lib.gpu_rng_streams(50, seed_state=raw.reshape(-1))
    assert np.array_equal(got, ol.prng(50, seed_state=raw.reshape(-1)))


def test_random_real_bit_exact(gpu_lib):
    rng = ol.prng(256, seed=1)
    out, _, _ = gpu_lib.gpu_sample("real", rng.copy(), 100, 0.0)
    for i in range(0, 256, 37):
        expect, _ = ol.random_real(rng[i], 100)
        assert np.array_equal(out[i], expect)


@pytest.mark.parametrize("kind,a,b", [
    ("binomial", [3, 20, 90, 1000, 5000, 1e6], [0.01, 0.3, 0.5, 0.07, 0.9, 0.002]),
    ("exponential", [0.5, 1.0, 1e6], None),
    ("normal", [0.0, 3.0], [1.0, 0.5]),
    ("poisson", [0.3, 4.0, 20.0, 800.0], None),
])
def test_samplers_fp64_draw_for_draw(gpu_lib, kind, a, b):
    n_streams, n_draws = 600, 60
    rng = ol.prng(n_streams, seed=77)
    aa = np.resize(np.array(a, dtype=float), n_streams)
    bb = None if b is None else np.resize(np.array(b, dtype=float), n_streams)
    out, err, rng_after = gpu_lib.gpu_sample(kind, rng.copy(), n_draws, aa, bb)
    assert not err.any()
    mismatch = 0
    for i in range(n_streams):
        expect, s_after = ol.sample(kind, rng[i], np.full(n_draws, aa[i]),
                                    None if bb is None else bb[i])
        same = np.array_equal(out[i], expect)
        if kind in ("exponential", "normal"):
            # closed-form transforms: device and host libm may differ in the
            # last place only
            np.testing.assert_allclose(out[i], expect, rtol=1e-14)
            same = True
        if same:
            assert np.array_equal(rng_after[i], s_after)
        else:
            # a last-place libm difference flipped an accept/reject decision
            mismatch += 1
    assert mismatch <= n_streams // 200, f"{mismatch} of {n_streams} streams diverged"


def chisq_pvalue(draws, pmf, support):
    from scipy import stats
    obs = np.array([(draws == k).sum() for k in support], dtype=float)
    exp = pmf(support) * draws.size
    keep = exp >= 5
    obs = np.append(obs[keep], draws.size - obs[keep].sum())
    exp = np.append(exp[keep], draws.size - exp[keep].sum())
    if exp[-1] < 5:
        obs[-2] += obs[-1]
        exp[-2] += exp[-1]
        obs, exp = obs[:-1], exp[:-1]
    return stats.chisquare(obs, exp * obs.sum() / exp.sum()).pvalue


@pytest.mark.parametrize("n,p", [(20, 0.1), (9, 0.5), (5, 0.93), (25, 0.4), (20, 0.5),
                                 (40, 0.3), (1000, 0.3), (60, 0.8), (1_000_000, 0.001),
                                 (200, 0.05), (1e6, 2e-5), (5e5, 0.5)])
def test_binomial_fp32_distribution(gpu_lib, n, p):
    """The FP32 sampler has no draw sequence to follow: its draws are held to
    the exact binomial pmf (chi-square) and mean.  Cases cover inversion
    (n p < 10), BTRS near its threshold where the Stirling remainders of k
    and n - k take their tabulated values, large n, and p > 1/2."""
    from scipy import stats
    rng = ol.prng(1000, seed=int(11 + n * p))
    out, err, _ = gpu_lib.gpu_sample("binomial", rng, 200, float(n), float(p),
                                     precision="fp32")
    assert not err.any()
    draws = out.reshape(-1)
    assert np.all(draws == np.round(draws))
    assert draws.min() >= 0 and draws.max() <= n
    mean = n * p
    sd = np.sqrt(n * p * (1 - p))
    lo, hi = int(max(0, mean - 8 * sd)), int(min(n, mean + 8 * sd)) + 1
    support = np.arange(lo, hi)
    assert chisq_pvalue(draws, lambda k: stats.binom.pmf(k, n, p), support) > 1e-4
    assert abs(draws.mean() - mean) < 5 * sd / np.sqrt(draws.size)


def test_binomial_fp32_mixed_wavefronts(gpu_lib):
    """Neighbouring lanes of one wavefront draw with different parameters:
    inversion (n q < 10) and BTRS lanes side by side, either side of the
    threshold, with p > 1/2 flipped and trivial lanes (n = 0, p = 0, p = 1)
    between them, so that lanes leave the search and rejection loops at
    different iterations.  Each parameter set's draws are held to its exact
    pmf."""
    from scipy import stats
    params = [(30, 0.2), (100, 0.3), (1e6, 9.9e-6), (1e6, 1.01e-5), (0, 0.3), (19, 0.52),
              (400, 0.95), (12, 1.0), (5000, 0.0), (21, 0.5), (2000, 0.01), (3, 0.4)]
    n_streams, n_draws = 64 * 60, 100
    nn = np.array([params[i % len(params)][0] for i in range(n_streams)], dtype=float)
    pp = np.array([params[i % len(params)][1] for i in range(n_streams)], dtype=float)
    rng = ol.prng(n_streams, seed=5)
    out, err, _ = gpu_lib.gpu_sample("binomial", rng, n_draws, nn, pp, precision="fp32")
    assert not err.any()
    for j, (n, p) in enumerate(params):
        draws = out[j::len(params)].reshape(-1)
        assert np.all(draws == np.round(draws)) and draws.min() >= 0 and draws.max() <= n
        if n == 0 or p == 0:
            assert np.all(draws == 0)
            continue
        if p == 1:
            assert np.all(draws == n)
            continue
        sd = np.sqrt(n * p * (1 - p))
        lo, hi = int(max(0, n * p - 8 * sd)), int(min(n, n * p + 8 * sd)) + 1
        support = np.arange(lo, hi)
        assert chisq_pvalue(draws, lambda k: stats.binom.pmf(k, n, p), support) > 1e-4, (n, p)
        assert abs(draws.mean() - n * p) < 5 * sd / np.sqrt(draws.size), (n, p)


def test_exponential_normal_poisson_fp32_distribution(gpu_lib):
    """The FP32 exponential (the compare noise of the FP32 filter), normal and
    Poisson samplers against their exact distributions, parameters
    interleaved lane by lane: Kolmogorov-Smirnov for the continuous ones,
    chi-square on the pmf for the Poisson."""
    from scipy import stats
    n_streams, n_draws = 64 * 40, 100
    rates = [0.5, 1.0, 1e6]
    rng = ol.prng(n_streams, seed=8)
    out, err, _ = gpu_lib.gpu_sample("exponential", rng, n_draws,
                                     np.resize(rates, n_streams), precision="fp32")
    assert not err.any()
    for j, rate in enumerate(rates):
        draws = out[j::len(rates)].reshape(-1)
        assert draws.min() >= 0
        assert stats.kstest(draws, stats.expon(scale=1 / rate).cdf).pvalue > 1e-4, rate
    mu_sd = [(0.0, 1.0), (3.0, 0.5), (-20.0, 4.0)]
    rng = ol.prng(n_streams, seed=9)
    out, err, _ = gpu_lib.gpu_sample("normal", rng, n_draws,
                                     np.resize([m for m, _ in mu_sd], n_streams),
                                     np.resize([s for _, s in mu_sd], n_streams), precision="fp32")
    assert not err.any()
    for j, (mu, sd) in enumerate(mu_sd):
        draws = out[j::len(mu_sd)].reshape(-1)
        assert stats.kstest(draws, stats.norm(mu, sd).cdf).pvalue > 1e-4, (mu, sd)
    means = [0.3, 4.0, 9.5, 20.0, 800.0]
    rng = ol.prng(n_streams, seed=10)
    out, err, _ = gpu_lib.gpu_sample("poisson", rng, n_draws, np.resize(means, n_streams),
                                     precision="fp32")
    assert not err.any()
    for j, lam in enumerate(means):
        draws = out[j::len(means)].reshape(-1)
        assert np.all(draws == np.round(draws)) and draws.min() >= 0
        sd = np.sqrt(lam)
        support = np.arange(int(max(0, lam - 8 * sd)), int(lam + 8 * sd) + 2)
        assert chisq_pvalue(draws, lambda k: stats.poisson.pmf(k, lam), support) > 1e-4, lam
        assert abs(draws.mean() - lam) < 5 * sd / np.sqrt(draws.size), lam


def test_binomial_invalid_reports(gpu_lib):
    rng = ol.prng(2, seed=1)
    out, err, _ = gpu_lib.gpu_sample("binomial", rng, 1, [10.0, 10.0], [0.5, -0.2])
    assert list(err) == [0, 1]


# ---------------------------------------------------------------------------
def weight_vectors():
    rng = np.random.default_rng(3)
    for n in [1, 2, 3, 10, 100, 511, 1000, 2048]:
        for scale in [0.1, 5, 300]:
            w = rng.normal(-40, scale, size=n)
            yield w, rng.random()
        w = rng.normal(-40, 5, size=n)
        if n > 2:
            w[rng.integers(0, n, n // 4 + 1)] = np.nan
            w[rng.integers(0, n, n // 4 + 1)] = -np.inf
        yield w, 0.0
        yield np.round(rng.normal(-40, 20, size=n) / 8) * 8, 1 - 1e-8
    yield np.full(16, -np.inf), 0.4
    yield np.full(64, -1.0), 0.0     # exact ties with every partial sum
    yield np.full(2048, -1.0), 0.0
    yield np.zeros(100), 0.5         # all zero: no resampling


def test_resample_bit_exact(gpu_lib):
    n_fallback = 0
    for w, u in weight_vectors():
        idx, out = gpu_lib.gpu_resample(w, u)
        if np.all(w == 0):
            assert out[1] == 0 and idx.tolist() == list(range(w.size))
            continue
        ll, scaled = ol.scale_log_weights(w)
        expect = ol.resample_weight(scaled, u) if w.size > 1 else np.zeros(1, np.uint32)
        assert idx.tolist() == expect.tolist(), (w.size, u)
        if np.isfinite(ll):
            assert out[0] == pytest.approx(ll, rel=1e-13, abs=1e-13)
        else:
            assert out[0] == ll
        n_fallback += int(out[2])
    # the tie cases must have been detected and resolved sequentially
    assert n_fallback >= 2


# ---------------------------------------------------------------------------
def run_gpu_filter(lib, pars, P, time, dt, data, seed, precision="fp64", launch=None,
                   save=False, engine="auto", graph=True):
    G = len(pars)
    obj = lib.dust_filter_create(lib.sir(), 0.0, {"time": time, "incidence": data}, P,
                                 n_groups=G, dt=dt, seed=seed, precision=precision)
    if launch:
        obj.set_launch(*launch)
    obj.set_engine(engine, graph=graph)
    plist = [dict(zip(["N", "I0", "beta", "gamma", "exp_noise"], p)) for p in pars]
    ll = lib.dust_likelihood_run(obj, plist, save_trajectories=save)
    return obj, np.atleast_1d(ll)


ENGINES = ["persistent", "pipeline"]


@pytest.mark.parametrize("engine", ENGINES)
@pytest.mark.parametrize("G,P,n_time", [(1, 1000, 100), (3, 100, 30), (2, 257, 40)])
def test_filter_fp64_matches_oracle(gpu_lib, G, P, n_time, engine):
    pars = np.array([[1000, 10, 0.2 + 0.05 * g, 0.1, 1e6] for g in range(G)])
    time = np.arange(1.0, n_time + 1)
    data = np.stack([sir_data(n_time, seed=42 + g) for g in range(G)], axis=1)
    obj, ll = run_gpu_filter(gpu_lib, pars, P, time, 0.25, data, 42, save=True, engine=engine)
    ref = ol.sir_filter(pars, P, 0.0, time, 0.25, data, seed=42, ancestry=True, n_threads=4)
    np.testing.assert_allclose(ll, ref["ll"], rtol=1e-12)
    state = gpu_lib.dust_likelihood_last_state(obj)
    assert np.array_equal(state, ref["state"])
    anc = gpu_lib.dust_likelihood_last_ancestry(obj)
    assert np.array_equal(anc, ref["ancestry"])
    rng = gpu_lib.dust_likelihood_rng_state(obj).reshape(-1, 4)
    assert np.array_equal(rng, ref["rng"])
    traj = gpu_lib.dust_likelihood_last_trajectories(obj)
    assert np.array_equal(traj[-1], ref["state"])


def test_pipeline_large_groups_match_oracle(gpu_lib):
    # more particles than one workgroup holds in registers
    pars = np.array([[5e5, 10, 0.4, 0.1, 1e6]])
    time = np.arange(1.0, 11.0)
    data = sir_data(10)
    obj, ll = run_gpu_filter(gpu_lib, pars, 10000, time, 0.25, data, 8, engine="pipeline")
    ref = ol.sir_filter(pars, 10000, 0.0, time, 0.25, data[:, None], seed=8, n_threads=8)
    np.testing.assert_allclose(ll, ref["ll"], rtol=1e-12)
    assert np.array_equal(gpu_lib.dust_likelihood_last_state(obj), ref["state"])


def test_pipeline_graph_replays_match_eager(gpu_lib):
    G, P = 4, 300
    pars = np.array([[2000, 10, 0.25 + 0.05 * g, 0.1, 1e6] for g in range(G)])
    time = np.arange(1.0, 31.0)
    data = sir_data(30)
    eager, ll_eager = run_gpu_filter(gpu_lib, pars, P, time, 0.25, data, 13, engine="pipeline",
                                     graph=False)
    rng0 = ol.prng(G * (P + 2), seed=13).reshape(-1)
    obj = gpu_lib.dust_filter_create(gpu_lib.sir(), 0.0, {"time": time, "incidence": data}, P,
                                     n_groups=G, dt=0.25, seed=13)
    obj.set_engine("pipeline", graph=True)
    plist = [dict(zip(["N", "I0", "beta", "gamma", "exp_noise"], p)) for p in pars]
    for rep in range(4):
        if rep:
            gpu_lib.dust_likelihood_set_rng_state(obj, rng0)
        # eager work on the device between replays
        gpu_lib.gpu_resample(np.random.default_rng(rep).normal(size=100), 0.5)
        ll = gpu_lib.dust_likelihood_run(obj, plist)
        assert np.array_equal(ll, ll_eager), rep
        assert np.array_equal(gpu_lib.dust_likelihood_last_state(obj),
                              gpu_lib.dust_likelihood_last_state(eager))


def test_filter_launch_shape_invariance(gpu_lib):
    pars = np.array([SIR, [1000, 10, 0.3, 0.1, 1e6]])
    time = np.arange(1.0, 21.0)
    data = sir_data(20)[:, None]
    results = []
    for launch in [(64, 4), (128, 2), (256, 1 * 2), (512, 2), (1024, 2)]:
        obj, ll = run_gpu_filter(gpu_lib, pars, 250, time, 0.25, data, 5, launch=launch)
        results.append((ll, gpu_lib.dust_likelihood_last_state(obj)))
    for ll, st in results[1:]:
        assert np.array_equal(ll, results[0][0])
        assert np.array_equal(st, results[0][1])


@pytest.mark.parametrize("engine", ENGINES)
def test_filter_missing_data_and_index_group(gpu_lib, engine):
    G, P = 3, 64
    pars = np.array([[1000, 10, 0.15 + 0.05 * g, 0.1, 1e6] for g in range(G)])
    time = np.arange(1.0, 26.0)
    data = np.stack([sir_data(25, seed=g) for g in range(G)], axis=1)
    data[[2, 3, 17], :] = np.nan
    data[9, 1] = np.nan
    obj, ll = run_gpu_filter(gpu_lib, pars, P, time, 0.25, data, 99, engine=engine)
    ref = ol.sir_filter(pars, P, 0.0, time, 0.25, data, seed=99)
    np.testing.assert_allclose(ll, ref["ll"], rtol=1e-12)
    # a subset of groups: selected groups reproduce the full run
    st = gpu_lib.dust_likelihood_rng_state(obj)
    obj2 = gpu_lib.dust_filter_create(gpu_lib.sir(), 0.0, {"time": time, "incidence": data}, P,
                                      n_groups=G, dt=0.25, seed=99)
    obj2.set_engine(engine)
    plist = [dict(zip(["N", "I0", "beta", "gamma", "exp_noise"], p)) for p in pars]
    # test-filter.R:39-51: an uninitialised object runs every group
    with pytest.raises(ValueError, match="'index_group' must be NULL, as 'obj' is not"):
        gpu_lib.dust_likelihood_run(obj2, plist[:1], index_group=[0])
    assert gpu_lib.dust_likelihood_ensure_initialised(obj2, plist)
    assert not gpu_lib.dust_likelihood_ensure_initialised(obj2, plist)
    ll_sub = gpu_lib.dust_likelihood_run(obj2, [plist[2], plist[0]], index_group=[2, 0])
    np.testing.assert_allclose(ll_sub, ll[[2, 0]], rtol=1e-12)
    with pytest.raises(ValueError, match="must be distinct"):
        gpu_lib.dust_likelihood_run(obj2, None, index_group=[1, 1])
    with pytest.raises(ValueError, match="must be at most 2"):
        gpu_lib.dust_likelihood_run(obj2, None, index_group=[3])
    del st


@pytest.mark.parametrize("engine", ENGINES)
def test_filter_errors_reported(gpu_lib, engine):
    pars = np.array([[1000, 10, 0.1, -0.2, 1e6]])
    time = np.arange(1.0, 4.0)
    data = np.array([1.0, 2.0, 3.0])
    with pytest.raises(gpu_lib.DustError) as e:
        run_gpu_filter(gpu_lib, pars, 10, time, 1.0, data, 42, engine=engine)
    msg = str(e.value)
    assert msg.startswith("10 particles reported errors.")
    assert "  - 1: Invalid call to binomial with" in msg
    assert "  - (and 6 more)" in msg


def test_filter_fp32_statistically_matches_fp64_oracle(gpu_lib):
    # 1000 replicate FP32 filters of the 1000-particle SIR problem (the
    # reference's example case) against 1000 replicates of the FP64 oracle,
    # each replicate on its own streams.  Bars: the mean log-likelihoods
    # within 1 combined Monte Carlo standard error (the bar stated for the
    # FP32 modes; the seeds are fixed, so the outcome is deterministic:
    # measured z = 0.03), the variance ratio inside the two-sided 99.9 %
    # F(999, 999) interval (measured 0.93 in [0.81, 1.23]) and the two
    # samples indistinguishable by a two-sample Kolmogorov-Smirnov test at
    # the 0.1 % level (measured p = 0.72).
    from scipy import stats
    G, P, n_time = 1000, 1000, 100
    time = np.arange(1.0, n_time + 1)
    data = sir_data(n_time)
    pars = np.array([SIR] * G)
    _, ll32 = run_gpu_filter(gpu_lib, pars, P, time, 0.25, data, 1, precision="fp32",
                             engine="pipeline")
    ref = ol.sir_filter(pars, P, 0.0, time, 0.25, data[:, None], seed=2, n_threads=16)
    ll64 = ref["ll"]
    se = np.sqrt(ll32.var(ddof=1) / ll32.size + ll64.var(ddof=1) / ll64.size)
    z = (ll32.mean() - ll64.mean()) / se
    assert abs(z) < 1, (ll32.mean(), ll64.mean(), se)
    ratio = ll32.var(ddof=1) / ll64.var(ddof=1)
    lo, hi = stats.f.ppf([0.0005, 0.9995], G - 1, G - 1)
    assert lo < ratio < hi, (ratio, lo, hi)
    assert stats.ks_2samp(ll32, ll64).pvalue > 1e-3


# ---------------------------------------------------------------------------
def test_system_simulate_draw_for_draw(gpu_lib):
    G, P = 2, 300
    pars = [dict(zip(["N", "I0", "beta", "gamma", "exp_noise"], p))
            for p in [[1000, 10, 1.0, 0.5, 1e6], [5e5, 50, 0.4, 0.1, 1e6]]]
    sys = gpu_lib.dust_system_create(gpu_lib.sir(), pars, P, n_groups=G, dt=0.25, seed=42)
    gpu_lib.dust_system_set_state_initial(sys)
    times = np.arange(0.0, 21.0)
    got = gpu_lib.dust_system_simulate(sys, times)
    rows = np.array([[p[k] for k in ["N", "I0", "beta", "gamma", "exp_noise"]] for p in pars])
    expect, rng_after = ol.sir_simulate(rows, P, 0.0, 0.25, times, ol.prng(G * P, seed=42))
    assert np.array_equal(got, expect)
    assert np.array_equal(gpu_lib.dust_system_rng_state(sys).reshape(-1, 4), rng_after)
    # incidence resets add up to cumulative cases
    np.testing.assert_array_equal(np.cumsum(got[:, :, :, 4], axis=0), got[:, :, :, 3])


def test_system_compare_and_reorder(gpu_lib):
    P = 10
    pars = {"beta": 0.1, "gamma": 0.2, "N": 1000, "I0": 10, "exp_noise": 0.5}
    sys = gpu_lib.dust_system_create(gpu_lib.sir(), pars, P, seed=42)
    s = np.zeros((1, P, 5))
    s[0, :, 4] = np.random.default_rng(1).poisson(30, P)
    gpu_lib.dust_system_set_state(sys, s)
    got = gpu_lib.dust_system_compare_data(sys, [30.0])
    # test-sir.R:31-47: Poisson(30; inc + Exp(0.5)) with the particle's own stream
    rng = ol.prng(P, seed=42)
    from scipy import stats
    for p in range(P):
        eps, _ = ol.sample("exponential", rng[p], [0.5])
        assert got[0, p] == pytest.approx(stats.poisson.logpmf(30, s[0, p, 4] + eps[0]),
                                          rel=1e-12)
    missing = gpu_lib.dust_system_compare_data(sys, [np.nan])
    assert np.all(missing == 0)
    idx = np.array([3, 3, 0, 9, 1, 2, 2, 5, 6, 7], dtype=np.uint32)
    gpu_lib.dust_system_reorder(sys, idx)
    np.testing.assert_array_equal(gpu_lib.dust_system_state(sys)[0], s[0][idx])


def test_system_errors_lock_until_reset(gpu_lib):
    pars = {"beta": 0.1, "gamma": -0.2, "N": 1000, "I0": 10, "exp_noise": 1e6}
    sys = gpu_lib.dust_system_create(gpu_lib.sir(), pars, 10, seed=42)
    gpu_lib.dust_system_set_state_initial(sys)
    with pytest.raises(gpu_lib.DustError, match="10 particles reported errors"):
        gpu_lib.dust_system_run_to_time(sys, 30)
    with pytest.raises(gpu_lib.DustError,
                       match="Can't currently run this system: errors are pending"):
        gpu_lib.dust_system_run_to_time(sys, 30)
    gpu_lib.dust_system_update_pars(sys, {"gamma": 0.2})
    gpu_lib.dust_system_set_state_initial(sys)
    assert gpu_lib.dust_system_time(sys) == 0


def test_autotune_keeps_results_and_streams(gpu_lib):
    G, P = 3, 500
    pars = np.array([[2000, 10, 0.25 + 0.05 * g, 0.1, 1e6] for g in range(G)])
    time = np.arange(1.0, 21.0)
    data = np.stack([sir_data(20, seed=7 + g) for g in range(G)], axis=1)
    plist = [dict(zip(["N", "I0", "beta", "gamma", "exp_noise"], p)) for p in pars]
    ref = gpu_lib.dust_filter_create(gpu_lib.sir(), 0.0, {"time": time, "incidence": data}, P,
                                     n_groups=G, dt=0.25, seed=11)
    ll_ref = gpu_lib.dust_likelihood_run(ref, plist)
    obj = gpu_lib.dust_filter_create(gpu_lib.sir(), 0.0, {"time": time, "incidence": data}, P,
                                     n_groups=G, dt=0.25, seed=11)
    report = obj.autotune(plist)
    assert len(report) >= 2 and all(r["ms"] > 0 for r in report)
    assert {r["engine"] for r in report} == {"persistent", "pipeline"}
    ll = gpu_lib.dust_likelihood_run(obj, None)
    np.testing.assert_array_equal(ll, ll_ref)
    np.testing.assert_array_equal(gpu_lib.dust_likelihood_rng_state(obj),
                                  gpu_lib.dust_likelihood_rng_state(ref))


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_results_do_not_depend_on_launch_shape(gpu_lib, precision):
    # the resampling indices are exact and the log-likelihood takes the
    # weights' aligned pairwise sum, so every engine, persistent launch
    # shape, pipeline chain count and resample workgroup width gives the
    # same bits: log-likelihoods, trajectories, final states and streams
    G, P, T = 4, 1000, 30
    pars = np.array([[2000 + 500 * g, 10, 0.25 + 0.05 * g, 0.1, 1e6] for g in range(G)])
    time = np.arange(1.0, T + 1)
    data = np.stack([sir_data(T, seed=21 + g) for g in range(G)], axis=1)
    plist = [dict(zip(["N", "I0", "beta", "gamma", "exp_noise"], p)) for p in pars]
    configs = [("pipeline", c, b) for c in (1, 2, 4) for b in (1024, 512, 256)]
    configs += [("persistent", 256, 4), ("persistent", 512, 2), ("persistent", 1024, 2)]
    out = []
    for engine, a, b in configs:
        obj = gpu_lib.dust_filter_create(gpu_lib.sir(), 0.0, {"time": time, "incidence": data},
                                         P, n_groups=G, dt=0.25, seed=5, precision=precision)
        if engine == "persistent":
            obj.set_launch(a, b)
            obj.set_engine("persistent")
        else:
            obj.set_engine("pipeline", chains=a, resample_block=b)
        ll = np.atleast_1d(gpu_lib.dust_likelihood_run(obj, plist, save_trajectories=True))
        assert gpu_lib.dust_likelihood_last_stats(obj)["engine"] == engine
        out.append((ll, gpu_lib.dust_likelihood_last_trajectories(obj),
                    gpu_lib.dust_likelihood_last_state(obj),
                    gpu_lib.dust_likelihood_rng_state(obj)))
    for cfg, r in zip(configs[1:], out[1:]):
        for a, b in zip(out[0], r):
            assert np.array_equal(a, b), cfg


NAMES = ["N", "I0", "beta", "gamma", "exp_noise"]


@pytest.mark.parametrize("engine", ENGINES)
def test_filter_initial_state_rng_state_and_parameter_change(gpu_lib, engine):
    # test-filter.R:207-307: manual initial state, parameters changed between
    # runs (the streams carry on), and streams set before a run
    G, P, T = 2, 300, 25
    pars = np.array([[2000, 10, 0.3, 0.1, 1e6], [5000, 20, 0.25, 0.12, 1e6]])
    time = np.arange(1.0, T + 1)
    data = np.stack([sir_data(T, seed=3 + g) for g in range(G)], axis=1)
    obj, ll1 = run_gpu_filter(gpu_lib, pars, P, time, 0.25, data, 9, engine=engine)
    rng1 = gpu_lib.dust_likelihood_rng_state(obj).reshape(-1, 4)
    ref1 = ol.sir_filter(pars, P, 0.0, time, 0.25, data, seed=9, n_threads=4)
    np.testing.assert_allclose(ll1, ref1["ll"], rtol=1e-12)
    # second run with new beta continues the streams
    pars2 = pars.copy()
    pars2[:, 2] = [0.35, 0.2]
    ll2 = gpu_lib.dust_likelihood_run(obj, [{"beta": b} for b in pars2[:, 2]])
    ref2 = ol.sir_filter(pars2, P, 0.0, time, 0.25, data, rng=rng1, n_threads=4)
    np.testing.assert_allclose(ll2, ref2["ll"], rtol=1e-12)
    # running does not touch filter stream 2g+1 (test-filter.R:118-148)
    rng2 = gpu_lib.dust_likelihood_rng_state(obj).reshape(G, P + 2, 4)
    assert np.array_equal(rng2[:, 1], rng1.reshape(G, P + 2, 4)[:, 1])
    # a manual initial state
    rng_init = np.random.default_rng(5)
    init = np.zeros((G, P, 5))
    init[:, :, 1] = rng_init.integers(5, 50, (G, P))
    init[:, :, 0] = pars2[:, None, 0] - init[:, :, 1]
    rng3 = gpu_lib.dust_likelihood_rng_state(obj).reshape(-1, 4)
    ll3 = gpu_lib.dust_likelihood_run(obj, None, initial=init)
    ref3 = ol.sir_filter(pars2, P, 0.0, time, 0.25, data, rng=rng3, initial=init, n_threads=4)
    np.testing.assert_allclose(ll3, ref3["ll"], rtol=1e-12)
    assert np.array_equal(gpu_lib.dust_likelihood_last_state(obj), ref3["state"])
    # streams set before a run (test-filter.R:249-276)
    gpu_lib.dust_likelihood_set_rng_state(obj, rng1)
    ll4 = gpu_lib.dust_likelihood_run(obj, None)
    np.testing.assert_array_equal(ll4, ll2)


def test_shared_data_equals_replicated_data(gpu_lib):
    # test-filter.R:747-772
    G, P, T = 3, 200, 20
    pars = np.array([[2000, 10, 0.3 + 0.02 * g, 0.1, 1e6] for g in range(G)])
    time = np.arange(1.0, T + 1)
    d = sir_data(T, seed=11)
    _, ll_shared = run_gpu_filter(gpu_lib, pars, P, time, 0.25, d, 4)
    _, ll_rep = run_gpu_filter(gpu_lib, pars, P, time, 0.25, np.tile(d[:, None], (1, G)), 4)
    np.testing.assert_array_equal(ll_shared, ll_rep)


# ---------------------------------------------------------------------------
# Random number state before and after runs (tests/testthat/test-filter.R)
def sir_pars(beta):
    return {"beta": beta, "gamma": 0.2, "N": 1000, "I0": 10, "exp_noise": 1e6}


SMALL = {"time": np.array([4.0, 8.0, 12.0, 16.0]), "incidence": np.arange(1.0, 5.0)}


def test_running_filter_keeps_second_filter_stream(gpu_lib):
    # test-filter.R:118-147
    P = 100
    obj = gpu_lib.dust_filter_create(gpu_lib.sir(), 0.0, SMALL, P, dt=1, seed=42)
    s0 = gpu_lib.dust_likelihood_rng_state(obj).reshape(-1, 4)
    assert s0.shape == (P + 2, 4)
    np.testing.assert_array_equal(s0, ol.prng(P + 2, seed=42))
    gpu_lib.dust_likelihood_run(obj, sir_pars(0.1))
    s1 = gpu_lib.dust_likelihood_rng_state(obj).reshape(-1, 4)
    np.testing.assert_array_equal(s1[1], s0[1])
    gpu_lib.dust_likelihood_run(obj, sir_pars(0.1))
    s2 = gpu_lib.dust_likelihood_rng_state(obj).reshape(-1, 4)
    np.testing.assert_array_equal(s2[1], s0[1])
    gpu_lib.dust_likelihood_last_state(obj, select_random_particle=True)
    s3 = gpu_lib.dust_likelihood_rng_state(obj).reshape(-1, 4)
    _, after = ol.random_real(s1[1], 1)
    np.testing.assert_array_equal(s3[1], after)
    np.testing.assert_array_equal(np.delete(s3, 1, axis=0), np.delete(s2, 1, axis=0))


def test_nested_filter_groups_equal_single_filters(gpu_lib):
    # test-filter.R:150-204: group g of a grouped filter is the filter whose
    # seed is the first stream of group g's block
    P = 100
    data = {"time": SMALL["time"], "incidence": np.column_stack([np.arange(1.0, 5.0),
                                                                 np.arange(2.0, 6.0)])}
    pars = [sir_pars(0.1), sir_pars(0.2)]
    obj = gpu_lib.dust_filter_create(gpu_lib.sir(), 0.0, data, P, n_groups=2, seed=42)
    s = gpu_lib.dust_likelihood_rng_state(obj).reshape(-1, 4)
    np.testing.assert_array_equal(s, ol.prng(2 * (P + 2), seed=42))
    res = np.array([gpu_lib.dust_likelihood_run(obj, pars if i == 0 else None)
                    for i in range(5)])
    for g in range(2):
        one = {"time": SMALL["time"], "incidence": data["incidence"][:, g]}
        seed = 42 if g == 0 else s[P + 2:P + 3]
        obj1 = gpu_lib.dust_filter_create(gpu_lib.sir(), 0.0, one, P, seed=seed)
        np.testing.assert_array_equal(gpu_lib.dust_likelihood_rng_state(obj1).reshape(-1, 4),
                                      s[g * (P + 2):(g + 1) * (P + 2)])
        res1 = np.array([gpu_lib.dust_likelihood_run(obj1, pars[g] if i == 0 else None)
                         for i in range(5)]).reshape(-1)
        np.testing.assert_array_equal(res1, res[:, g])


@pytest.mark.parametrize("after_run", [False, True])
def test_set_rng_state_before_and_after_running(gpu_lib, after_run):
    # test-filter.R:249-307
    data = {"time": SMALL["time"], "incidence": np.column_stack([np.arange(1.0, 5.0),
                                                                 np.arange(2.0, 6.0)])}
    pars = [sir_pars(0.1), sir_pars(0.2)]
    obj1 = gpu_lib.dust_filter_create(gpu_lib.sir(), 0.0, data, 100, n_groups=2, seed=42)
    obj2 = gpu_lib.dust_filter_create(gpu_lib.sir(), 0.0, data, 100, n_groups=2, seed=43)
    if after_run:
        assert not np.array_equal(gpu_lib.dust_likelihood_run(obj1, pars),
                                  gpu_lib.dust_likelihood_run(obj2, pars))
    s1 = gpu_lib.dust_likelihood_rng_state(obj1)
    s2 = gpu_lib.dust_likelihood_rng_state(obj2)
    assert not np.array_equal(s1, s2)
    gpu_lib.dust_likelihood_set_rng_state(obj2, s1)
    np.testing.assert_array_equal(gpu_lib.dust_likelihood_rng_state(obj2), s1)
    np.testing.assert_array_equal(gpu_lib.dust_likelihood_run(obj1, pars),
                                  gpu_lib.dust_likelihood_run(obj2, pars))


def test_copy_is_independent(gpu_lib):
    # test-filter.R:310-340
    obj1 = gpu_lib.dust_filter_create(gpu_lib.sir(), 0.0, SMALL, 100, seed=42)
    obj2 = gpu_lib.dust_likelihood_copy(obj1)
    obj3 = gpu_lib.dust_likelihood_copy(obj1, seed=42)
    s1 = gpu_lib.dust_likelihood_rng_state(obj1)
    assert not np.array_equal(gpu_lib.dust_likelihood_rng_state(obj2), s1)
    np.testing.assert_array_equal(gpu_lib.dust_likelihood_rng_state(obj3), s1)
    ll1 = gpu_lib.dust_likelihood_run(obj1, sir_pars(0.1))
    ll2 = gpu_lib.dust_likelihood_run(obj2, sir_pars(0.1))
    ll3 = gpu_lib.dust_likelihood_run(obj3, sir_pars(0.1))
    assert ll1 == ll3 and ll1 != ll2


def test_partial_trajectories_by_index_state(gpu_lib):
    # test-filter.R:92-116 and :360-395
    pars = [sir_pars(0.1), sir_pars(0.2)]
    data = {"time": SMALL["time"], "incidence": np.arange(1.0, 9.0).reshape(2, 4).T}
    obj1 = gpu_lib.dust_filter_create(gpu_lib.sir(), 0.0, data, 100, n_groups=2, seed=42)
    obj2 = gpu_lib.dust_filter_create(gpu_lib.sir(), 0.0, data, 100, n_groups=2, seed=42)
    ll1 = gpu_lib.dust_likelihood_run(obj1, pars, save_trajectories=True)
    h1 = gpu_lib.dust_likelihood_last_trajectories(obj1)
    ll2 = gpu_lib.dust_likelihood_run(obj2, pars, save_trajectories=True, index_state=[1, 3])
    h2 = gpu_lib.dust_likelihood_last_trajectories(obj2)
    np.testing.assert_array_equal(ll1, ll2)
    assert h1.shape == (4, 2, 100, 5) and h2.shape == (4, 2, 100, 2)
    np.testing.assert_array_equal(h2, h1[..., [1, 3]])
    # per group, with a state subset
    obj3 = gpu_lib.dust_filter_create(gpu_lib.sir(), 0.0, data, 100, n_groups=2, seed=42)
    obj4 = gpu_lib.dust_filter_create(gpu_lib.sir(), 0.0, data, 100, n_groups=2, seed=42)
    assert gpu_lib.dust_likelihood_ensure_initialised(obj3, pars)
    assert gpu_lib.dust_likelihood_ensure_initialised(obj4, pars)
    la = gpu_lib.dust_likelihood_run(obj3, pars, index_state=[0, 2, 4], save_trajectories=True)
    ha = gpu_lib.dust_likelihood_last_trajectories(obj3)
    for g in range(2):
        lb = gpu_lib.dust_likelihood_run(obj4, [pars[g]], index_state=[0, 2, 4], index_group=[g],
                                         save_trajectories=True)
        np.testing.assert_array_equal(lb, la[[g]])
        np.testing.assert_array_equal(gpu_lib.dust_likelihood_last_trajectories(obj4),
                                      ha[:, [g]])
    # snapshots and the final state keep every state
    gpu_lib.dust_likelihood_run(obj1, None, save_snapshots=[8.0], index_state=[4])
    assert gpu_lib.dust_likelihood_last_snapshots(obj1).shape == (1, 2, 100, 5)
    assert gpu_lib.dust_likelihood_last_state(obj1).shape == (2, 100, 5)
    with pytest.raises(gpu_lib.DustError, match="must be increasing"):
        gpu_lib.dust_likelihood_run(obj1, None, save_snapshots=[12.0, 4.0])
    with pytest.raises(ValueError, match="continuous-time models"):
        gpu_lib.dust_filter_create(gpu_lib.sirode(), 0.0, SMALL, 10)


def test_pipeline_launch_chains_give_identical_results(gpu_lib):
    # the groups split into 1-4 launch chains on their own streams: same
    # results, and errors in groups of different chains are reported from
    # the earliest interval (dust2/discrete/system.hpp:87)
    G, P = 7, 200
    pars = np.array([[2000, 10, 0.2 + 0.04 * g, 0.1, 1e6] for g in range(G)])
    time = np.arange(1.0, 31.0)
    data = sir_data(30)
    out = []
    for chains in [1, 2, 3, 4]:
        obj = gpu_lib.dust_filter_create(gpu_lib.sir(), 0.0, {"time": time, "incidence": data},
                                         P, n_groups=G, dt=0.25, seed=21)
        obj.set_engine("pipeline", chains=chains)
        plist = [dict(zip(["N", "I0", "beta", "gamma", "exp_noise"], p)) for p in pars]
        ll = gpu_lib.dust_likelihood_run(obj, plist, save_trajectories=True)
        out.append((ll, gpu_lib.dust_likelihood_last_state(obj),
                    gpu_lib.dust_likelihood_last_trajectories(obj),
                    gpu_lib.dust_likelihood_rng_state(obj)))
        # groups 1 and 6 (different chains for 2-4 chains) fail; only the
        # particles of the earliest failing interval are reported
        bad = [dict(p) for p in plist]
        bad[1]["gamma"] = -0.5
        bad[6]["gamma"] = -0.5
        with pytest.raises(gpu_lib.DustError) as e:
            gpu_lib.dust_likelihood_run(obj, bad)
        assert str(e.value).startswith(f"{2 * P} particles reported errors."), chains
    for ll, st, tr, rng in out[1:]:
        np.testing.assert_array_equal(ll, out[0][0])
        np.testing.assert_array_equal(st, out[0][1])
        np.testing.assert_array_equal(tr, out[0][2])
        np.testing.assert_array_equal(rng, out[0][3])
    with pytest.raises(gpu_lib.DustError, match="between 1 and 4"):
        obj.set_engine("pipeline", chains=5)
