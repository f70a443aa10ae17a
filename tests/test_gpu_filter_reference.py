"""The reference's particle filter tests (tests/testthat/test-filter.R) not
covered elsewhere, on the GPU: pars = NULL and index_group on an
uninitialised filter, groups whose missing data fall at different times
equal separate filters seeded from their stream blocks, rows of missing
data from integer data equal removed rows, the final state against the
trajectories (with and without index_state, grouped), random particles
from a nested filter, and the snapshot argument checks."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

PARS = {"beta": 0.1, "gamma": 0.2, "N": 1000.0, "I0": 10.0, "exp_noise": 1e6}
DATA = {"time": np.array([4.0, 8.0, 12.0, 16.0]), "incidence": np.arange(1.0, 5.0)}


def test_pars_null_only_on_initialised_filter(gpu_lib):
    # test-filter.R:20-36
    obj1 = gpu_lib.dust_filter_create(gpu_lib.sir(), 0, DATA, 100, seed=42)
    obj2 = gpu_lib.dust_filter_create(gpu_lib.sir(), 0, DATA, 100, seed=42)
    with pytest.raises(ValueError, match="'pars' cannot be NULL, as 'obj' is not initialised"):
        gpu_lib.dust_likelihood_run(obj1, None)
    assert gpu_lib.dust_likelihood_run(obj1, PARS) == gpu_lib.dust_likelihood_run(obj2, PARS)
    assert gpu_lib.dust_likelihood_run(obj2, None) == gpu_lib.dust_likelihood_run(obj1, PARS)
    # test-filter.R:39-51
    data = {"time": DATA["time"],
            "incidence": np.column_stack([np.arange(1.0, 5.0), np.arange(2.0, 6.0)])}
    obj = gpu_lib.dust_filter_create(gpu_lib.sir(), 0, data, 10, n_groups=2)
    with pytest.raises(ValueError, match="'index_group' must be NULL, as 'obj' is not initialised"):
        gpu_lib.dust_likelihood_run(obj, [PARS], index_group=[0])


def test_groups_with_missing_data_at_different_times(gpu_lib):
    # test-filter.R:459-493: each group equals a filter of its own data,
    # seeded with its block of the grouped filter's streams
    pars = [PARS, dict(PARS, beta=0.2, gamma=0.1)]
    time = np.array([4.0, 8.0, 12.0, 16.0, 20.0])
    inc = np.array([[1, 2], [2, np.nan], [np.nan, 4], [np.nan, np.nan], [5, 6]], dtype=float)
    P = 100
    obj = gpu_lib.dust_filter_create(gpu_lib.sir(), 0, {"time": time, "incidence": inc}, P,
                                     n_groups=2, seed=42)
    streams = gpu_lib.dust_likelihood_rng_state(obj).reshape(2, P + 2, 4)
    obj1 = gpu_lib.dust_filter_create(gpu_lib.sir(), 0, {"time": time, "incidence": inc[:, 0]},
                                      P, seed=42)
    obj2 = gpu_lib.dust_filter_create(gpu_lib.sir(), 0, {"time": time, "incidence": inc[:, 1]},
                                      P, seed=streams[1].copy())
    ll = np.array([gpu_lib.dust_likelihood_run(obj, pars) for _ in range(10)])
    ll1 = [gpu_lib.dust_likelihood_run(obj1, pars[0]) for _ in range(10)]
    ll2 = [gpu_lib.dust_likelihood_run(obj2, pars[1]) for _ in range(10)]
    np.testing.assert_array_equal(ll[:, 0], ll1)
    np.testing.assert_array_equal(ll[:, 1], ll2)


def test_missing_rows_from_integer_data(gpu_lib):
    # test-filter.R:496-514: a missing first row equals the row removed
    data1 = {"time": DATA["time"], "incidence": np.array([np.nan, 2, 3, 4])}
    data2 = {"time": DATA["time"][1:], "incidence": np.array([2, 3, 4], dtype=np.int64)}
    obj1 = gpu_lib.dust_filter_create(gpu_lib.sir(), 0, data1, 100, seed=42)
    obj2 = gpu_lib.dust_filter_create(gpu_lib.sir(), 0, data2, 100, seed=42)
    ll1 = [gpu_lib.dust_likelihood_run(obj1, PARS) for _ in range(10)]
    ll2 = [gpu_lib.dust_likelihood_run(obj2, PARS) for _ in range(10)]
    assert ll1 == ll2


def test_final_state_against_trajectories(gpu_lib):
    # test-filter.R:570-628
    obj = gpu_lib.dust_filter_create(gpu_lib.sir(), 0, DATA, 10)
    gpu_lib.dust_likelihood_run(obj, PARS, save_trajectories=True)
    h = gpu_lib.dust_likelihood_last_trajectories(obj)
    s = gpu_lib.dust_likelihood_last_state(obj)
    assert s.shape == (1, 10, 5)
    np.testing.assert_array_equal(s, h[-1])
    gpu_lib.dust_likelihood_run(obj, PARS, save_trajectories=False)
    with pytest.raises(gpu_lib.DustError, match="Trajectories are not current"):
        gpu_lib.dust_likelihood_last_trajectories(obj)
    gpu_lib.dust_likelihood_last_state(obj)
    # index_state keeps some states in the trajectories, all in the state
    gpu_lib.dust_likelihood_run(obj, PARS, index_state=[0, 1, 2], save_trajectories=True)
    h = gpu_lib.dust_likelihood_last_trajectories(obj)
    s = gpu_lib.dust_likelihood_last_state(obj)
    np.testing.assert_array_equal(s[..., :3], h[-1])
    assert s.shape == (1, 10, 5)
    # grouped, with data per group
    pars = [dict(PARS, beta=b) for b in (0.1, 0.2, 0.3)]
    data = {"time": DATA["time"], "incidence": np.arange(1.0, 13.0).reshape(3, 4).T}
    obj = gpu_lib.dust_filter_create(gpu_lib.sir(), 0, data, 10, n_groups=3)
    gpu_lib.dust_likelihood_run(obj, pars, save_trajectories=True)
    np.testing.assert_array_equal(gpu_lib.dust_likelihood_last_state(obj),
                                  gpu_lib.dust_likelihood_last_trajectories(obj)[-1])


def test_random_particle_from_nested_filter(gpu_lib):
    # test-filter.R:667-690: each group's chosen particle is one of its
    # final particles, and not the same slot in every group
    pars = [dict(PARS, beta=b) for b in (0.1, 0.2, 0.3)]
    data = {"time": DATA["time"], "incidence": np.arange(1.0, 13.0).reshape(3, 4).T}
    obj = gpu_lib.dust_filter_create(gpu_lib.sir(), 0, data, 100, n_groups=3, seed=42)
    gpu_lib.dust_likelihood_run(obj, pars, save_trajectories=True)
    s1 = gpu_lib.dust_likelihood_last_state(obj)
    s2 = gpu_lib.dust_likelihood_last_state(obj, select_random_particle=True)
    assert s2.shape == (3, 5)
    slots = []
    for g in range(3):
        match = np.flatnonzero(np.all(s1[g] == s2[g], axis=1))
        assert match.size > 0
        slots.append(set(match.tolist()))
    assert not set.intersection(*slots) or len(set.union(*slots)) > 1


def test_snapshot_arguments_are_checked(gpu_lib):
    # test-filter.R:860-879
    obj = gpu_lib.dust_filter_create(gpu_lib.sir(), 0, DATA, 100, seed=42)
    with pytest.raises(Exception, match="Values in 'save_snapshots' must be increasing"):
        gpu_lib.dust_likelihood_run(obj, PARS, save_snapshots=[12.0, 4.0])
    for bad in ([4.1, 12.0], [4.1, 13.0]):
        with pytest.raises(Exception,
                           match="All elements of 'save_snapshots' must be found in 'time'"):
            gpu_lib.dust_likelihood_run(obj, PARS, save_snapshots=bad)
