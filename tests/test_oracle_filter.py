"""Pins the oracle's filter orchestration with the reference's invariant
tests (tests/testthat/test-filter.R, test-sir.R), restated against the
oracle: grouped filters equal independent filters seeded from the expanded
stream array, missing data equals removed data, results do not depend on
the thread count, incidence resets sum to cumulative cases."""
import numpy as np
import pytest

import oracle_lib as ol

SIR = [1000.0, 10.0, 0.2, 0.1, 1e6]


def sir_data(n_time=30, seed=1):
    rng = np.random.default_rng(seed)
    return rng.poisson(5, size=n_time).astype(np.float64)


def test_grouped_equals_independent():
    # test-filter.R:151-204
    G, P = 3, 20
    pars = np.array([[1000, 10, 0.15 + 0.05 * g, 0.1, 1e6] for g in range(G)])
    time = np.arange(1.0, 21.0)
    data = np.stack([sir_data(20, g) for g in range(G)], axis=1)
    res = ol.sir_filter(pars, P, 0.0, time, 0.25, data, seed=42)
    streams = ol.prng(G * (P + 2), seed=42)
    for g in range(G):
        one = ol.sir_filter(pars[g:g + 1], P, 0.0, time, 0.25, data[:, g:g + 1],
                            rng=streams[g * (P + 2):(g + 1) * (P + 2)])
        assert one["ll"][0] == res["ll"][g]
        assert np.array_equal(one["state"][0], res["state"][g])


def test_missing_data_equals_removed_rows():
    # test-filter.R:439-514
    P = 50
    time = np.arange(1.0, 31.0)
    d = sir_data(30)
    d_na = d.copy()
    d_na[[3, 4, 10, 29]] = np.nan
    keep = ~np.isnan(d_na)
    a = ol.sir_filter(np.array([SIR]), P, 0.0, time, 0.25, d_na[:, None], seed=7)
    b = ol.sir_filter(np.array([SIR]), P, 0.0, time[keep], 0.25, d[keep][:, None], seed=7)
    assert a["ll"][0] == b["ll"][0]


def test_thread_count_invariance():
    G, P = 4, 64
    pars = np.array([[1000, 10, 0.2 + 0.02 * g, 0.1, 1e6] for g in range(G)])
    time = np.arange(1.0, 16.0)
    data = sir_data(15)[:, None]
    a = ol.sir_filter(pars, P, 0.0, time, 0.25, data, seed=3, n_threads=1)
    b = ol.sir_filter(pars, P, 0.0, time, 0.25, data, seed=3, n_threads=4)
    assert np.array_equal(a["ll"], b["ll"])
    assert np.array_equal(a["state"], b["state"])
    assert np.array_equal(a["rng"], b["rng"])


def test_rng_state_roundtrip_reproduces_run():
    # test-filter.R:249-307: same rng state => same likelihood
    P = 30
    time = np.arange(1.0, 11.0)
    data = sir_data(10)[:, None]
    a = ol.sir_filter(np.array([SIR]), P, 0.0, time, 0.25, data, seed=11)
    b = ol.sir_filter(np.array([SIR]), P, 0.0, time, 0.25, data, rng=a["rng"])
    c = ol.sir_filter(np.array([SIR]), P, 0.0, time, 0.25, data, rng=a["rng"])
    assert b["ll"][0] == c["ll"][0]
    assert b["ll"][0] != a["ll"][0]


def test_single_particle_likelihood_is_sum_of_log_weights():
    # scale_log_weights returns the raw weight when P == 1 (filter_details.hpp:13-15)
    time = np.arange(1.0, 6.0)
    data = sir_data(5)[:, None]
    res = ol.sir_filter(np.array([SIR]), 1, 0.0, time, 1.0, data, seed=5)
    assert np.isfinite(res["ll"][0]) and res["ll"][0] < 0


def test_incidence_resets_sum_to_cumulative_cases():
    # test-sir.R:98-135
    pars = np.array([[1000, 10, 1.0, 0.5, 1e6]])
    P = 100
    rng = ol.prng(P, seed=42)
    res, _ = ol.sir_simulate(pars, P, 0.0, 0.25, np.arange(0.0, 6.0), rng)
    inc = res[:, 0, :, 4]
    cum = res[:, 0, :, 3]
    assert np.all(np.diff(cum, axis=0) >= 0)
    assert np.any(np.diff(inc, axis=0) < 0)
    np.testing.assert_array_equal(np.cumsum(inc, axis=0), cum)


def test_invalid_parameters_report_errors():
    # test-sir.R:251-287
    pars = np.array([[1000, 10, 0.1, -0.2, 1e6]])
    with pytest.raises(ol.OracleError) as e:
        ol.sir_filter(pars, 10, 0.0, np.arange(1.0, 4.0), 1.0, sir_data(3)[:, None], seed=42)
    msg = str(e.value)
    assert msg.startswith("10 particles reported errors.")
    assert "  - 1: Invalid call to binomial with" in msg
    assert "  - (and 6 more)" in msg
