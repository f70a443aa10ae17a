"""The particle filter over a mixed-time model on the GPU (the malaria
model; tests/testthat/test-filter.R:705-722): against the whole filter run
through the reference solver and the reference's own resampling
(tests/golden/malaria_filter.json: log-likelihoods of consecutive runs
within 1e-10, final states within 1e-9), and against a filter assembled by
hand from the system's entry points as the reference's test does
(tests/testthat/helper-dust.R:5-48, filter_manual)."""
import json
import os

import numpy as np
import pytest

import oracle_lib as ol

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "malaria_filter.json")


def golden():
    with open(GOLDEN) as f:
        return json.load(f)["cases"]


def data_dict(case):
    d = np.array([[[np.nan if v is None else v for v in f] for f in g] for g in case["data"]],
                 dtype=float)
    return {"time": np.array(case["times"]), "tested": d[:, :, 0], "positive": d[:, :, 1]}


@pytest.mark.parametrize("case", golden(), ids=lambda c: c["name"])
def test_matches_reference_filter(gpu_lib, case):
    G, P = len(case["pars"]), case["P"]
    streams = np.array(case["streams"], dtype=np.uint64)
    obj = gpu_lib.dust_filter_create(gpu_lib.malaria(), 0.0, data_dict(case), P, n_groups=G,
                                     dt=case["dt"] or None, seed=streams)
    pars = case["pars"] if G > 1 else case["pars"][0]
    for r in case["runs"]:
        ll = np.atleast_1d(gpu_lib.dust_likelihood_run(obj, pars))
        np.testing.assert_allclose(ll, r["ll"], rtol=1e-10)
        st = gpu_lib.dust_likelihood_last_state(obj)
        np.testing.assert_allclose(st, np.array(r["state"]), rtol=1e-9, atol=1e-12)
    # the particle streams never move (updates draw from copies); the
    # filter streams advanced once per resampling
    rng = gpu_lib.dust_likelihood_rng_state(obj).reshape(G, P + 2, 4)
    np.testing.assert_array_equal(rng[:, 2:], streams.reshape(G, P + 2, 4)[:, 2:])


def test_filter_equals_filter_by_hand(gpu_lib):
    # test-filter.R:705-722 with filter_manual (helper-dust.R:5-48)
    times = np.array([4.0, 8.0, 12.0, 16.0])
    data = {"time": times, "tested": [2, 4, 6, 8], "positive": [1, 2, 3, 3]}
    P = 100
    streams = ol.prng(P + 2, seed=42)
    obj = gpu_lib.dust_filter_create(gpu_lib.malaria(), 0.0, data, P, dt=1, seed=streams)
    res = [gpu_lib.dust_likelihood_run(obj, {}) for _ in range(5)]

    sys = gpu_lib.dust_system_create(gpu_lib.malaria(), {}, P, dt=1, seed=streams[2:])
    fr = streams[0].copy()
    manual = []
    for _ in range(5):
        gpu_lib.dust_system_set_time(sys, 0.0)
        gpu_lib.dust_system_set_state_initial(sys)
        ll = 0.0
        for i, t in enumerate(times):
            gpu_lib.dust_system_run_to_time(sys, t)
            tmp = gpu_lib.dust_system_compare_data(
                sys, {"tested": data["tested"][i], "positive": data["positive"][i]})[0]
            inc, w = ol.scale_log_weights(tmp)
            ll += inc
            u, fr = ol.random_real(fr, 1)
            k = ol.resample_weight(w, u[0])
            state = gpu_lib.dust_system_state(sys)
            gpu_lib.dust_system_set_state(sys, state[:, k])
        manual.append(ll)
    np.testing.assert_allclose(res, manual, rtol=1e-12)


def test_trajectories_follow_the_lineage(gpu_lib):
    case = golden()[0]
    P = case["P"]
    obj = gpu_lib.dust_filter_create(gpu_lib.malaria(), 0.0, data_dict(case), P, dt=1,
                                     seed=np.array(case["streams"], dtype=np.uint64))
    with pytest.raises(gpu_lib.DustError, match="Trajectories are not current"):
        gpu_lib.dust_likelihood_last_trajectories(obj)
    ll = gpu_lib.dust_likelihood_run(obj, {}, save_trajectories=True)
    assert ll == pytest.approx(case["runs"][0]["ll"][0], rel=1e-10)
    h = gpu_lib.dust_likelihood_last_trajectories(obj)
    assert h.shape == (4, 1, P, 20)
    np.testing.assert_array_equal(h[-1, 0], gpu_lib.dust_likelihood_last_state(obj)[0])
    # each particle's path is continuous in the never-resampled compartment
    # totals: humans Sh + Ih stay 1 along every lineage
    np.testing.assert_allclose(h[..., 0] + h[..., 1], 1.0, rtol=1e-9)


def test_select_random_particle(gpu_lib):
    # filter<T>::select_random_particle (dust2/filter.hpp:130-138): one draw
    # per group of the last run from its second filter stream, chosen
    # particle floor(u * P), the same particle for the trajectories,
    # snapshots and state, kept until the next run (:173)
    times = np.array([3.0, 6.0, 9.0])
    G, P = 3, 40
    data = {"time": times, "tested": np.full((3, G), 6.0),
            "positive": np.array([[1, 2, 3], [2, 3, 1], [3, 1, 2]], dtype=float)}
    pars = [{"a": 0.3}, {"a": 0.35}, {"a": 0.4}]
    obj = gpu_lib.dust_filter_create(gpu_lib.malaria(), 0.0, data, P, n_groups=G, dt=1,
                                     seed=ol.prng(G * (P + 2), seed=11))
    gpu_lib.dust_likelihood_run(obj, pars, save_trajectories=True, save_snapshots=[6.0, 9.0])
    s1 = gpu_lib.dust_likelihood_rng_state(obj).reshape(G, P + 2, 4)
    h = gpu_lib.dust_likelihood_last_trajectories(obj)
    snap = gpu_lib.dust_likelihood_last_snapshots(obj)
    st = gpu_lib.dust_likelihood_last_state(obj)
    h1 = gpu_lib.dust_likelihood_last_trajectories(obj, select_random_particle=True)
    assert h1.shape == (3, G, 1, 20)
    s2 = gpu_lib.dust_likelihood_rng_state(obj).reshape(G, P + 2, 4)
    for g in range(G):
        u, after = ol.random_real(s1[g, 1], 1)
        assert np.array_equal(s2[g, 1], after)
        k = int(np.floor(u[0] * P))
        np.testing.assert_array_equal(h1[:, g, 0], h[:, g, k])
        np.testing.assert_array_equal(
            gpu_lib.dust_likelihood_last_snapshots(obj, select_random_particle=True)[:, g, 0],
            snap[:, g, k])
        np.testing.assert_array_equal(
            gpu_lib.dust_likelihood_last_state(obj, select_random_particle=True)[g], st[g, k])
    assert np.array_equal(np.delete(s2, 1, axis=1), np.delete(s1, 1, axis=1))
    # cached: no further draws
    assert np.array_equal(gpu_lib.dust_likelihood_rng_state(obj).reshape(G, P + 2, 4), s2)
    # a run over a subset: the choice is redrawn for the groups of that run
    gpu_lib.dust_likelihood_run(obj, None, index_group=[2, 0], save_trajectories=True)
    s3 = gpu_lib.dust_likelihood_rng_state(obj).reshape(G, P + 2, 4)
    h = gpu_lib.dust_likelihood_last_trajectories(obj)
    h1 = gpu_lib.dust_likelihood_last_trajectories(obj, select_random_particle=True)
    assert h1.shape == (3, 2, 1, 20)
    for i, g in enumerate([2, 0]):
        u, _ = ol.random_real(s3[g, 1], 1)
        np.testing.assert_array_equal(h1[:, i, 0], h[:, i, int(np.floor(u[0] * P))])
    s4 = gpu_lib.dust_likelihood_rng_state(obj).reshape(G, P + 2, 4)
    assert np.array_equal(s4[1], s3[1])
    # deterministic unfilters take their first particle
    un = gpu_lib.dust_unfilter_create(gpu_lib.malaria(), 0.0, data, n_particles=2, n_groups=G)
    gpu_lib.dust_likelihood_run(un, pars)
    np.testing.assert_array_equal(
        gpu_lib.dust_likelihood_last_state(un, select_random_particle=True),
        gpu_lib.dust_likelihood_last_state(un)[:, 0])


def test_filter_errors(gpu_lib):
    data = {"time": [4.0, 8.0], "tested": [2, 4], "positive": [1, 2]}
    with pytest.raises(ValueError,
                       match=r"'time_start' \(5\) is later than the first time in 'data' \(4\)"):
        gpu_lib.dust_filter_create(gpu_lib.malaria(), 5.0, data, 10, dt=1)
    with pytest.raises(ValueError, match="continuous-time models"):
        gpu_lib.dust_filter_create(gpu_lib.sirode(), 0.0, {"time": [1.0], "incidence": [1.0]},
                                   10)


def test_unfilter_equals_the_reference_solver_run(gpu_lib):
    # dust_unfilter_create on the malaria model (dust2/unfilter.hpp:42-72):
    # deterministic, so the likelihood is the binomial log density of the
    # data at the reference solver's trajectory (the malaria_continuous
    # golden case, a = 0.4, no updates)
    from scipy import stats
    with open(os.path.join(os.path.dirname(__file__), "golden", "ode_mixed_dp5.json")) as f:
        case = next(c for c in json.load(f)["cases"] if c["name"] == "malaria_continuous")
    times = np.array(case["targets"][1:])
    ih = np.array(case["particles"][0]["state"])[1:, 1]
    tested = np.arange(times.size) % 7 + 3.0
    positive = np.floor(tested * 0.3)
    positive[4] = np.nan
    data = {"time": times, "tested": tested, "positive": positive}
    obj = gpu_lib.dust_unfilter_create(gpu_lib.malaria(), 0.0, data, n_particles=3)
    ll = gpu_lib.dust_likelihood_run(obj, {"a": 0.4})
    ok = ~np.isnan(positive)
    want = stats.binom.logpmf(positive[ok], tested[ok], ih[ok]).sum()
    np.testing.assert_allclose(ll, want, rtol=1e-10)
    # deterministic particles are identical; trajectories are the states at
    # the data times
    assert np.all(ll == ll[0, 0])
    gpu_lib.dust_likelihood_run(obj, None, save_trajectories=True)
    h = gpu_lib.dust_likelihood_last_trajectories(obj)
    np.testing.assert_allclose(h[:, 0, 0, 1], ih, rtol=1e-10)
    np.testing.assert_array_equal(gpu_lib.dust_likelihood_last_state(obj), h[-1])
    with pytest.raises(ValueError, match="only available for particle filters"):
        gpu_lib.dust_likelihood_rng_state(obj)
    # with dt the deterministic updates keep beta, so only the solver's
    # stopping points change
    obj1 = gpu_lib.dust_unfilter_create(gpu_lib.malaria(), 0.0, data, dt=1)
    np.testing.assert_allclose(gpu_lib.dust_likelihood_run(obj1, {"a": 0.4}), want, rtol=1e-6)


def test_index_group_runs_a_subset(gpu_lib):
    # R/interface-likelihood.R:54-75: a subset of groups equals those groups
    # of the whole run, their streams untouched by the others
    times = np.array([3.0, 6.0, 9.0])
    G, P = 3, 32
    data = {"time": times, "tested": np.full((3, G), 6.0),
            "positive": np.array([[1, 2, 3], [2, 3, 1], [3, 1, 2]], dtype=float)}
    pars = [{"a": 0.3}, {"a": 0.35}, {"a": 0.4}]
    streams = ol.prng(G * (P + 2), seed=7)
    whole = gpu_lib.dust_filter_create(gpu_lib.malaria(), 0.0, data, P, n_groups=G, dt=1,
                                       seed=streams)
    ll = gpu_lib.dust_likelihood_run(whole, pars, save_trajectories=True)
    h = gpu_lib.dust_likelihood_last_trajectories(whole)
    part = gpu_lib.dust_filter_create(gpu_lib.malaria(), 0.0, data, P, n_groups=G, dt=1,
                                      seed=streams)
    gpu_lib.dust_likelihood_run(part, pars)
    gpu_lib.dust_likelihood_set_rng_state(part, streams)
    sub = gpu_lib.dust_likelihood_run(part, None, index_group=[2, 0], save_trajectories=True)
    np.testing.assert_array_equal(sub, ll[[2, 0]])
    np.testing.assert_array_equal(gpu_lib.dust_likelihood_last_trajectories(part), h[:, [2, 0]])
    # the group left out keeps its initial state and its streams
    st = gpu_lib.dust_likelihood_last_state(part)
    np.testing.assert_array_equal(st[1], np.broadcast_to(st[1, :1], st[1].shape))
    rng = gpu_lib.dust_likelihood_rng_state(part).reshape(G, P + 2, 4)
    np.testing.assert_array_equal(rng[1], streams.reshape(G, P + 2, 4)[1])
    # the unfilter: a subset's rows equal the whole run's
    un = gpu_lib.dust_unfilter_create(gpu_lib.malaria(), 0.0, data, n_particles=2, n_groups=G)
    ll_un = gpu_lib.dust_likelihood_run(un, pars)
    np.testing.assert_array_equal(gpu_lib.dust_likelihood_run(un, None, index_group=[1]),
                                  ll_un[[1]])
    with pytest.raises(ValueError, match="'index_group' must be NULL"):
        gpu_lib.dust_likelihood_run(
            gpu_lib.dust_filter_create(gpu_lib.malaria(), 0.0, data, P, n_groups=G, dt=1),
            pars, index_group=[0])
    obj = gpu_lib.dust_filter_create(gpu_lib.malaria(), 0.0, data, 10, n_groups=G, dt=1)
    gpu_lib.dust_likelihood_run(obj, pars)
    for fn in [gpu_lib.dust_likelihood_last_stats, gpu_lib.dust_likelihood_last_ancestry]:
        with pytest.raises(ValueError, match="not available for filters over continuous-time"):
            fn(obj)
