"""The reference's SIR system tests (tests/testthat/test-sir.R) not covered
elsewhere, on the GPU: creation and the initial state, compare_data with
missing data (no draw) and across parameter groups (the exponential noise
from each particle's own stream, checked through the oracle's sampler),
the data length check, running to a time with dt, parameter updates,
simulate with and without a state index, reorder with replacement, the
RNG state and setting the state from one vector."""
import numpy as np
import pytest
from scipy import stats

import oracle_lib as ol

pytestmark = pytest.mark.gpu

PARS = {"beta": 0.1, "gamma": 0.2, "N": 1000.0, "I0": 10.0, "exp_noise": 1e6}


def test_simple_sir_system(gpu_lib):
    # test-sir.R:1-28
    obj = gpu_lib.dust_system_create(gpu_lib.sir(), PARS, 10, seed=42)
    assert gpu_lib.sir().n_state == 5
    assert gpu_lib.dust_system_rng_state(obj).nbytes == 32 * 10
    np.testing.assert_array_equal(gpu_lib.dust_system_state(obj), np.zeros((1, 10, 5)))
    assert gpu_lib.dust_system_time(obj) == 0
    gpu_lib.dust_system_set_state_initial(obj)
    np.testing.assert_array_equal(gpu_lib.dust_system_state(obj)[0],
                                  np.tile([990, 10, 0, 0, 0], (10, 1)))
    gpu_lib.dust_system_run_to_time(obj, 30)
    s1 = gpu_lib.dust_system_state(obj)[0]
    assert np.all(s1[:, 0] < 990) and np.all(s1[:, 2] > 0) and np.all(s1[:, 3] > 0)


def test_compare_missing_data_draws_nothing(gpu_lib):
    # test-sir.R:48-61
    obj = gpu_lib.dust_system_create(gpu_lib.sir(), dict(PARS, exp_noise=0.5), 10, seed=42)
    s = np.zeros((1, 10, 5))
    s[0, :, 4] = np.random.default_rng(1).poisson(30, 10)
    gpu_lib.dust_system_set_state(obj, s)
    before = gpu_lib.dust_system_rng_state(obj)
    np.testing.assert_array_equal(gpu_lib.dust_system_compare_data(obj, [np.nan]), 0)
    np.testing.assert_array_equal(gpu_lib.dust_system_rng_state(obj), before)
    np.testing.assert_array_equal(before, ol.prng(10, seed=42).reshape(-1))


def test_compare_across_parameter_groups(gpu_lib):
    # test-sir.R:64-84: each particle adds Exp(exp_noise) noise drawn from its
    # own stream before the Poisson density
    pars = [dict(PARS, beta=0.1 * i, exp_noise=10.0 ** i) for i in range(1, 5)]
    obj = gpu_lib.dust_system_create(gpu_lib.sir(), pars, 10, n_groups=4, seed=42)
    s = gpu_lib.dust_system_state(obj)
    s[:, :, 4] = np.random.default_rng(2).poisson(30, (4, 10))
    gpu_lib.dust_system_set_state(obj, s)
    streams = gpu_lib.dust_system_rng_state(obj).reshape(40, 4)
    res = gpu_lib.dust_system_compare_data(obj, [31.0, 32.0, 33.0, 34.0])
    want = np.zeros((4, 10))
    for g in range(4):
        for p in range(10):
            eps, _ = ol.sample("exponential", streams[g * 10 + p], [10.0 ** (g + 1)])
            want[g, p] = stats.poisson.logpmf(31 + g, s[g, p, 4] + eps[0])
    np.testing.assert_allclose(res, want, rtol=1e-12)
    # test-sir.R:87-95
    with pytest.raises(ValueError, match="Expected 'data' to have length 4, but it had length 3"):
        gpu_lib.dust_system_compare_data(obj, [1.0, 2.0, 3.0])


def test_run_to_time_update_pars_and_simulate(gpu_lib):
    # test-sir.R:138-154: dt = 0.25, the same seed gives the same state
    a = gpu_lib.dust_system_create(gpu_lib.sir(), PARS, 10, dt=0.25, seed=42)
    b = gpu_lib.dust_system_create(gpu_lib.sir(), PARS, 10, dt=0.25, seed=42)
    for x in (a, b):
        gpu_lib.dust_system_set_state_initial(x)
        gpu_lib.dust_system_run_to_time(x, 10)
        assert gpu_lib.dust_system_time(x) == 10
    np.testing.assert_array_equal(gpu_lib.dust_system_state(a), gpu_lib.dust_system_state(b))
    # test-sir.R:157-172: update_pars equals creating with the combined set
    a = gpu_lib.dust_system_create(gpu_lib.sir(), PARS, 10, seed=42)
    gpu_lib.dust_system_update_pars(a, {"beta": 0.3, "gamma": 0.2})
    gpu_lib.dust_system_run_to_time(a, 10)
    b = gpu_lib.dust_system_create(gpu_lib.sir(), dict(PARS, beta=0.3), 10, seed=42)
    gpu_lib.dust_system_run_to_time(b, 10)
    np.testing.assert_array_equal(gpu_lib.dust_system_state(a), gpu_lib.dust_system_state(b))
    # test-sir.R:175-206: simulate, with and without a state index
    a = gpu_lib.dust_system_create(gpu_lib.sir(), PARS, 10, seed=42)
    b = gpu_lib.dust_system_create(gpu_lib.sir(), PARS, 10, seed=42)
    c = gpu_lib.dust_system_create(gpu_lib.sir(), PARS, 10, seed=42)
    for x in (a, b, c):
        gpu_lib.dust_system_set_state_initial(x)
    res = gpu_lib.dust_system_simulate(a, np.arange(0.0, 21.0))
    assert res.shape == (21, 1, 10, 5)
    np.testing.assert_array_equal(res[-1], gpu_lib.dust_system_state(a))
    assert gpu_lib.dust_system_time(a) == 20
    np.testing.assert_array_equal(res[0], gpu_lib.dust_system_state(b))
    gpu_lib.dust_system_run_to_time(b, 10)
    np.testing.assert_array_equal(res[10], gpu_lib.dust_system_state(b))
    np.testing.assert_array_equal(
        gpu_lib.dust_system_simulate(c, np.arange(0.0, 21.0), index_state=[1, 3]),
        res[..., [1, 3]])


def test_reorder_rng_state_and_vector_state(gpu_lib):
    # test-sir.R:225-234: reorder with replacement
    obj = gpu_lib.dust_system_create(gpu_lib.sir(), {}, 10, seed=42)
    gpu_lib.dust_system_set_state_initial(obj)
    gpu_lib.dust_system_run_to_time(obj, 100)
    s1 = gpu_lib.dust_system_state(obj)
    i = np.random.default_rng(3).integers(0, 10, 10)
    gpu_lib.dust_system_reorder(obj, i)
    np.testing.assert_array_equal(gpu_lib.dust_system_state(obj), s1[:, i])
    # test-sir.R:237-248
    a = gpu_lib.dust_system_create(gpu_lib.sir(), PARS, 10, seed=42)
    b = gpu_lib.dust_system_create(gpu_lib.sir(), PARS, 10, seed=43)
    assert not np.array_equal(gpu_lib.dust_system_rng_state(a), gpu_lib.dust_system_rng_state(b))
    gpu_lib.dust_system_set_rng_state(b, gpu_lib.dust_system_rng_state(a))
    np.testing.assert_array_equal(gpu_lib.dust_system_rng_state(a),
                                  gpu_lib.dust_system_rng_state(b))
    # test-sir.R:301-308
    sys = gpu_lib.dust_system_create(gpu_lib.sir(), PARS, 10)
    gpu_lib.dust_system_set_state(sys, [1000, 10, 0, 0, 0])
    np.testing.assert_array_equal(gpu_lib.dust_system_state(sys)[0],
                                  np.tile([1000, 10, 0, 0, 0], (10, 1)))
