"""The reference's unfilter tests (tests/testthat/test-unfilter.R) on the
GPU: the likelihood computed by hand from the deterministic system, groups
against separate unfilters, replicated particles, subsets of groups and
states, the final state, and the continuous-time unfilter against the
continuous-time system."""
import numpy as np
import pytest
from scipy import stats

pytestmark = pytest.mark.gpu

BASE = {"beta": 0.1, "gamma": 0.2, "N": 1000.0, "I0": 10.0, "exp_noise": 1e6}
DATA = {"time": np.array([4.0, 8.0, 12.0, 16.0]), "incidence": np.arange(1.0, 5.0)}


def grouped_data(n_groups):
    return {"time": DATA["time"],
            "incidence": np.arange(1.0, 4 * n_groups + 1).reshape(n_groups, 4).T}


def manual_ll(lib, pars, initial=None, gen="sir"):
    # test-unfilter.R:1-31: run the deterministic system to each data time
    full = dict(BASE, **pars)
    if gen == "sir":
        obj = lib.dust_system_create(lib.sir(), full, 1, time=0, dt=1, deterministic=True)
        noise = 1e-6
    else:
        full = {k: full[k] for k in ["N", "I0", "beta", "gamma"]}
        obj = lib.dust_system_create(lib.sirode(), full, 1, time=0, deterministic=True)
        noise = 0
    if initial is None:
        lib.dust_system_set_state_initial(obj)
    else:
        lib.dust_system_set_state(obj, initial)
    inc = []
    for t in DATA["time"]:
        lib.dust_system_run_to_time(obj, t)
        inc.append(lib.dust_system_state(obj)[0, 0, 4])
    return stats.poisson.logpmf(DATA["incidence"], np.array(inc) + noise).sum()


def test_can_run_an_unfilter(gpu_lib):
    obj = gpu_lib.dust_unfilter_create(gpu_lib.sir(), 0.0, DATA)
    p1 = {"beta": 0.1, "gamma": 0.2, "I0": 10.0}
    p2 = {"beta": 0.2, "gamma": 0.2, "I0": 10.0}
    assert gpu_lib.dust_likelihood_run(obj, BASE) == pytest.approx(manual_ll(gpu_lib, p1),
                                                                     rel=1e-12)
    assert gpu_lib.dust_likelihood_run(obj, p1) == pytest.approx(manual_ll(gpu_lib, p1),
                                                                   rel=1e-12)
    assert gpu_lib.dust_likelihood_run(obj, p2) == pytest.approx(manual_ll(gpu_lib, p2),
                                                                   rel=1e-12)
    # test-unfilter.R:33-43
    obj2 = gpu_lib.dust_unfilter_create(gpu_lib.sir(), 0.0, DATA)
    with pytest.raises(ValueError, match="'pars' cannot be NULL, as 'obj' is not initialised"):
        gpu_lib.dust_likelihood_run(obj2, None)
    ll = gpu_lib.dust_likelihood_run(obj2, BASE)
    assert gpu_lib.dust_likelihood_run(obj2, None) == ll


def test_manually_set_state(gpu_lib):
    # test-unfilter.R:96-121
    state = np.array([1000 - 17, 17, 0, 0, 0.0])
    obj = gpu_lib.dust_unfilter_create(gpu_lib.sir(), 0.0, DATA)
    ll = gpu_lib.dust_likelihood_run(obj, BASE, initial=state)
    assert ll == pytest.approx(manual_ll(gpu_lib, {}, initial=state), rel=1e-12)


def test_structured_and_replicated(gpu_lib):
    # test-unfilter.R:123-176
    pars = [dict(BASE, beta=0.1 * (i + 1)) for i in range(3)]
    data = grouped_data(3)
    obj = gpu_lib.dust_unfilter_create(gpu_lib.sir(), 0.0, data, n_groups=3)
    ll = gpu_lib.dust_likelihood_run(obj, pars)
    sep = [gpu_lib.dust_likelihood_run(
        gpu_lib.dust_unfilter_create(gpu_lib.sir(), 0.0,
                                     {"time": data["time"], "incidence": data["incidence"][:, g]}),
        pars[g]) for g in range(3)]
    np.testing.assert_allclose(ll, sep, rtol=1e-14)
    rep = gpu_lib.dust_unfilter_create(gpu_lib.sir(), 0.0, DATA, n_particles=5)
    one = gpu_lib.dust_unfilter_create(gpu_lib.sir(), 0.0, DATA, n_particles=1)
    np.testing.assert_array_equal(gpu_lib.dust_likelihood_run(rep, BASE),
                                  np.full((1, 5), gpu_lib.dust_likelihood_run(one, BASE)))
    rep2 = gpu_lib.dust_unfilter_create(gpu_lib.sir(), 0.0, grouped_data(2), n_particles=5,
                                        n_groups=2)
    one2 = gpu_lib.dust_unfilter_create(gpu_lib.sir(), 0.0, grouped_data(2), n_groups=2)
    p2 = [BASE, dict(BASE, beta=0.2)]
    np.testing.assert_array_equal(gpu_lib.dust_likelihood_run(rep2, p2),
                                  np.repeat(gpu_lib.dust_likelihood_run(one2, p2)[:, None], 5,
                                            axis=1))


def test_trajectories_subsets_and_final_state(gpu_lib):
    # test-unfilter.R:46-95, 250-347, 363-417
    obj = gpu_lib.dust_unfilter_create(gpu_lib.sir(), 0.0, DATA)
    with pytest.raises(gpu_lib.DustError, match="Trajectories are not current"):
        gpu_lib.dust_likelihood_last_trajectories(obj)
    gpu_lib.dust_likelihood_run(obj, BASE)
    with pytest.raises(gpu_lib.DustError, match="Trajectories are not current"):
        gpu_lib.dust_likelihood_last_trajectories(obj)
    gpu_lib.dust_likelihood_run(obj, None, save_trajectories=True)
    h = gpu_lib.dust_likelihood_last_trajectories(obj)
    m = gpu_lib.dust_system_create(gpu_lib.sir(), BASE, 1, time=0, deterministic=True)
    gpu_lib.dust_system_set_state_initial(m)
    np.testing.assert_array_equal(h, gpu_lib.dust_system_simulate(m, DATA["time"]))
    np.testing.assert_array_equal(gpu_lib.dust_likelihood_last_state(obj), h[-1])
    gpu_lib.dust_likelihood_run(obj, None, save_trajectories=False)
    with pytest.raises(gpu_lib.DustError, match="Trajectories are not current"):
        gpu_lib.dust_likelihood_last_trajectories(obj)
    np.testing.assert_array_equal(gpu_lib.dust_likelihood_last_state(obj), h[-1])
    for fn in [gpu_lib.dust_likelihood_rng_state, gpu_lib.dust_likelihood_last_stats,
               gpu_lib.dust_likelihood_last_ancestry]:
        with pytest.raises(ValueError, match="only available for particle filters"):
            fn(obj)
    with pytest.raises(gpu_lib.DustError, match="Snapshots are not current"):
        gpu_lib.dust_likelihood_last_snapshots(obj)
    # the C ABI refuses a handle of another type
    from dust2_amd._lib import call, dptr
    with pytest.raises(gpu_lib.DustError, match="Not a filter handle"):
        call("dust2_filter_last_state", obj.ptr, dptr(np.zeros(5)))
    # index_state: trajectories keep the chosen states, the final state all
    gpu_lib.dust_likelihood_run(obj, None, save_trajectories=True, index_state=[1, 3])
    np.testing.assert_array_equal(gpu_lib.dust_likelihood_last_trajectories(obj),
                                  h[..., [1, 3]])
    assert gpu_lib.dust_likelihood_last_state(obj).shape == (1, 1, 5)
    # groups: a subset of an unfilter reproduces the whole run's groups
    pars0 = [{"beta": 0.01, "gamma": 0.01, "N": 100.0, "I0": 1.0, "exp_noise": 1e6}] * 3
    pars1 = [{"beta": 0.1 * (i + 1), "gamma": 0.2, "I0": 10.0, "exp_noise": 1e6}
             for i in range(3)]
    data = grouped_data(3)
    o1 = gpu_lib.dust_unfilter_create(gpu_lib.sir(), 0.0, data, n_groups=3)
    gpu_lib.dust_likelihood_run(o1, pars0)
    ll1 = gpu_lib.dust_likelihood_run(o1, pars1, save_trajectories=True)
    h1 = gpu_lib.dust_likelihood_last_trajectories(o1)
    o2 = gpu_lib.dust_unfilter_create(gpu_lib.sir(), 0.0, data, n_groups=3)
    gpu_lib.dust_likelihood_run(o2, pars0)
    np.testing.assert_array_equal(
        gpu_lib.dust_likelihood_run(o2, pars1[::-1], index_group=[2, 1, 0],
                                    save_trajectories=True), ll1[::-1])
    np.testing.assert_array_equal(gpu_lib.dust_likelihood_last_trajectories(o2),
                                  h1[:, [2, 1, 0]])
    for i in range(3):
        gpu_lib.dust_likelihood_run(o2, [pars1[i]], index_group=[i], save_trajectories=True,
                                    index_state=[0, 2, 4])
        np.testing.assert_array_equal(gpu_lib.dust_likelihood_last_trajectories(o2),
                                      h1[:, [i]][..., [0, 2, 4]])


def test_continuous_time_unfilter(gpu_lib):
    # test-unfilter.R:419-446
    obj = gpu_lib.dust_unfilter_create(gpu_lib.sirode(), 0.0, DATA)
    p1 = {"beta": 0.1, "gamma": 0.2, "I0": 10.0}
    p2 = {"beta": 0.2, "gamma": 0.2, "I0": 10.0}
    full = {k: BASE[k] for k in ["N", "I0", "beta", "gamma"]}
    assert gpu_lib.dust_likelihood_run(obj, full) == pytest.approx(
        manual_ll(gpu_lib, p1, gen="sirode"), rel=1e-6)
    assert gpu_lib.dust_likelihood_run(obj, p2) == pytest.approx(
        manual_ll(gpu_lib, p2, gen="sirode"), rel=1e-6)
