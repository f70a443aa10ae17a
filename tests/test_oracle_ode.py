"""Pins the oracle's Dormand-Prince 5(4) unfilter against the reference's own
dust2/continuous/solver.hpp (golden vectors from oracle/_ref: daily states,
step counters, final step size, log-likelihood) and the logistic closed form
(tests/testthat/helper-dust.R:68-73)."""
import json
import os

import numpy as np
import pytest

import oracle_lib as ol

HERE = os.path.join(os.path.dirname(__file__), "golden")


def load(name):
    with open(os.path.join(HERE, name)) as f:
        return json.load(f)


def test_sirode_unfilter_matches_reference_solver():
    g = load("sirode_dp5.json")
    pars = np.array(g["pars"])
    res = ol.sirode_unfilter(pars, 1, g["time_start"], g["time"], np.array(g["data"]))
    # bit-identical: same algorithm, same operation order, same libm
    assert res["ll"].tolist() == g["ll"]
    assert res["stats"].tolist() == g["stats"]
    final = np.array(g["state"])[:, -1, :]
    assert np.array_equal(res["state"][:, 0, :], final)


def test_sirode_tight_control_matches_reference_solver():
    g = load("sirode_dp5.json")
    pars = np.array(g["pars"])[:4]
    res = ol.sirode_unfilter(pars, 1, 0.0, g["time"][:60], np.array(g["data"][:60]),
                             control=g["control2"])
    assert res["ll"].tolist() == g["ll2"]
    assert res["stats"].tolist() == g["stats2"]
    assert np.array_equal(res["state"][:, 0, :], np.array(g["state2_final"]))


def test_logistic_closed_form_through_reference():
    g = load("logistic_dp5.json")
    t = np.array(g["time"])
    y = np.array(g["y"])
    K, y0 = g["K"], g["y0"]
    for j, r in enumerate([g["r1"], g["r2"]]):
        exact = K / (1 + (K / y0 - 1) * np.exp(-r * t))
        np.testing.assert_allclose(y[:, j], exact, rtol=1e-5)


def test_unfilter_deterministic_sir_closed_form():
    # test-unfilter.R:1-30: deterministic SIR ll = sum dpois(obs, inc + 1e-6)
    from scipy import stats
    pars = np.array([[1000, 10, 0.1, 0.2, 1e6], [1000, 10, 0.2, 0.2, 1e6]], dtype=float)
    time = np.array([4.0, 8.0, 12.0, 16.0])
    data = np.tile(np.arange(1.0, 5.0)[:, None], (1, 2))
    ll, _ = ol.sir_unfilter(pars, 0.0, time, 1.0, data, gradient=False)
    for g in range(2):
        N, I0, beta, gamma = pars[g, :4]
        S, I, R, cum = N - I0, I0, 0.0, 0.0
        inc = 0.0
        expect = 0.0
        t = 0
        for i, tt in enumerate(time):
            while t < tt:
                inc = 0.0 if t == int(t) else inc  # daily reset at integer times
                p_SI = 1 - np.exp(-beta * I / N)
                p_IR = 1 - np.exp(-gamma)
                n_SI, n_IR = S * p_SI, I * p_IR
                S, I, R = S - n_SI, I + n_SI - n_IR, R + n_IR
                inc += n_SI
                t += 1
            expect += stats.poisson.logpmf(data[i, g], inc + 1e-6)
        assert ll[g] == pytest.approx(expect, rel=1e-12)


def _fd_gradient(pars, time, dt, data, rel=1e-6):
    # central differences in (beta, gamma, I0), N held fixed
    out = np.zeros((pars.shape[0], 3))
    for j, col in enumerate([2, 3, 1]):
        h = rel * pars[:, col]
        up, dn = pars.copy(), pars.copy()
        up[:, col] += h
        dn[:, col] -= h
        lu, _ = ol.sir_unfilter(up, 0.0, time, dt, data, gradient=False)
        ld, _ = ol.sir_unfilter(dn, 0.0, time, dt, data, gradient=False)
        out[:, j] = (lu - ld) / (2 * h)
    return out


@pytest.mark.parametrize("times,dt", [([4.0, 8.0, 12.0, 16.0], 1.0),    # test-unfilter-adjoint.R:14-27
                                      ([3.0, 6.0, 9.0, 12.0], 1.0),     # odd step counts, :120-133
                                      ([4.0, 8.0, 12.0, 16.0], 0.25)])
def test_adjoint_gradient_matches_finite_differences(times, dt):
    pars = np.array([[1000, 10, 0.1, 0.2, 1e6], [5000, 25, 0.35, 0.12, 1e6],
                     [1000, 10, 0.3, 0.05, 1e6]], dtype=float)
    time = np.array(times)
    data = np.tile(np.arange(1.0, 5.0)[:, None], (1, 3))
    ll, grad = ol.sir_unfilter(pars, 0.0, time, dt, data, exact=True)
    ll2, _ = ol.sir_unfilter(pars, 0.0, time, dt, data, gradient=False)
    assert np.array_equal(ll, ll2)  # the adjoint run gives the same likelihood
    fd = _fd_gradient(pars, time, dt, data)
    np.testing.assert_allclose(grad, fd, rtol=1e-5)


def test_adjoint_gradient_with_missing_data_and_long_series():
    rng = np.random.default_rng(3)
    pars = np.array([[1e4, 10, 0.25, 0.1, 1e6]])
    time = np.arange(1.0, 61.0)
    data = rng.poisson(20, (60, 1)).astype(float)
    data[[3, 17, 40], 0] = np.nan
    _, grad = ol.sir_unfilter(pars, 0.0, time, 0.25, data, exact=True)
    fd = _fd_gradient(pars, time, 0.25, data)
    np.testing.assert_allclose(grad, fd, rtol=1e-5)


def _numpy_sir_ll(row, state0, time, dt, data):
    """Deterministic SIR log-likelihood from a given initial state, written
    from inst/examples/sir.cpp (binomial -> n p, Exp(rate) -> 1 / rate,
    incidence reset at integer times)."""
    from scipy import stats
    N, _, beta, gamma, noise = row
    S, I, R, cum, inc = state0
    t, ll = 0.0, 0.0
    for i, target in enumerate(time):
        n = int(round((target - t) / dt))
        for s in range(n):
            if np.fmod(t + s * dt, 1.0) == 0:
                inc = 0.0
            n_SI = S * (1 - np.exp(-beta * I / N * dt))
            n_IR = I * (1 - np.exp(-gamma * dt))
            S, I, R, cum, inc = S - n_SI, I + n_SI - n_IR, R + n_IR, cum + n_SI, inc + n_SI
        t += n * dt
        if not np.isnan(data[i]):
            ll += stats.poisson.logpmf(data[i], inc + 1 / noise)
    return ll


@pytest.mark.parametrize("times,dt", [([4.0, 8.0, 12.0, 16.0], 1.0),    # test-unfilter-adjoint.R:14-27
                                      ([3.0, 6.0, 9.0, 12.0], 1.0),     # :120-133
                                      ([4.0, 8.0, 12.0, 16.0], 0.25),
                                      ([2.5, 5.0, 7.5], 0.5)])
def test_reference_adjoint_pinned_by_what_it_differentiates(times, dt):
    # The reference's hand-written adjoint (inst/examples/sir.cpp:126-198)
    # differentiates: beta exactly; gamma through exp(-gamma) where p_IR's
    # derivative is dt exp(-gamma dt), a constant factor per group, so its
    # gamma gradient is the true one times exp(-gamma) / (dt exp(-gamma dt));
    # and I0 through I(0) alone, holding S(0) fixed.  Each is checked
    # against central differences of an independent numpy recursion.
    pars = np.array([[1000, 10, 0.1, 0.2, 1e6], [5000, 25, 0.35, 0.12, 1e6]], dtype=float)
    time = np.array(times)
    data = np.tile(np.arange(1.0, time.size + 1)[:, None], (1, 2))
    _, grad = ol.sir_unfilter(pars, 0.0, time, dt, data)
    fd = _fd_gradient(pars, time, dt, data)
    for g in range(2):
        N, I0, beta, gamma = pars[g, :4]
        assert grad[g, 0] == pytest.approx(fd[g, 0], rel=1e-5)
        factor = np.exp(-gamma) / (dt * np.exp(-gamma * dt))
        assert grad[g, 1] == pytest.approx(fd[g, 1] * factor, rel=1e-5)
        h = 1e-6 * I0
        up = _numpy_sir_ll(pars[g], [N - I0, I0 + h, 0, 0, 0], time, dt, data[:, g])
        dn = _numpy_sir_ll(pars[g], [N - I0, I0 - h, 0, 0, 0], time, dt, data[:, g])
        assert grad[g, 2] == pytest.approx((up - dn) / (2 * h), rel=1e-5)
        # the exact I0 gradient also moves S(0) = N - I0
        assert abs(grad[g, 2] - fd[g, 2]) > 1e-3 * abs(fd[g, 2])


def test_numpy_recursion_equals_oracle_likelihood():
    pars = np.array([[5000, 25, 0.35, 0.12, 1e6]])
    time = np.array([2.5, 5.0, 7.5])
    data = np.array([[3.0], [np.nan], [9.0]])
    ll, _ = ol.sir_unfilter(pars, 0.0, time, 0.5, data, gradient=False)
    N, I0 = pars[0, :2]
    assert ll[0] == pytest.approx(_numpy_sir_ll(pars[0], [N - I0, I0, 0, 0, 0], time, 0.5,
                                                data[:, 0]), rel=1e-12)


def test_reference_adjoint_is_nan_with_missing_data():
    # the reference's compare adjoint has no missing-data branch
    # (sir.cpp:179-197); the exact form skips missing data
    pars = np.array([[1e4, 10, 0.25, 0.1, 1e6]])
    time = np.arange(1.0, 11.0)
    data = np.full((10, 1), 5.0)
    data[4, 0] = np.nan
    _, ref = ol.sir_unfilter(pars, 0.0, time, 1.0, data)
    _, exact = ol.sir_unfilter(pars, 0.0, time, 1.0, data, exact=True)
    assert np.all(np.isnan(ref)) and np.all(np.isfinite(exact))


def _ode_fd(pars, time, data, control=None, rel=1e-6):
    out = np.zeros((pars.shape[0], 3))
    for j, col in enumerate([2, 3, 1]):
        h = rel * pars[:, col]
        up, dn = pars.copy(), pars.copy()
        up[:, col] += h
        dn[:, col] -= h
        lu, _ = ol.sirode_gradient(up, 0.0, time, data, control=control)
        ld, _ = ol.sirode_gradient(dn, 0.0, time, data, control=control)
        out[:, j] = (lu - ld) / (2 * h)
    return out


@pytest.mark.parametrize("control", [None, [100000, 1e-10, 1e-10, 0, np.inf]])
def test_ode_sensitivity_gradient_matches_finite_differences(control):
    rng = np.random.default_rng(1)
    pars = np.column_stack([np.full(6, 1000.0), rng.uniform(5, 20, 6), rng.uniform(0.1, 0.5, 6),
                            rng.uniform(0.05, 0.2, 6)])
    time = np.arange(1.0, 91.0)
    data = rng.poisson(8, (90, 6)).astype(float)
    data[[3, 40], :] = np.nan
    ll, grad = ol.sirode_gradient(pars, 0.0, time, data, control=control)
    # the sensitivity run integrates the same steps: the likelihood is the
    # plain unfilter's, bit for bit
    ctl = None if control is None else control
    ref = ol.sirode_unfilter(pars, 1, 0.0, time, data, control=ctl)["ll"]
    assert np.array_equal(ll, ref)
    np.testing.assert_allclose(grad, _ode_fd(pars, time, data, control), rtol=1e-6)


def test_sirode_weekly_data_matches_reference_solver():
    # steps spanning several daily resets: the reset subtracts the dense
    # output at the boundary, and after the last step the final correction
    g = load("sirode_weekly_dp5.json")
    pars = np.array(g["pars"])
    for ll_key, stats_key, state_key, ctl in [("ll", "stats", "state", None),
                                              ("ll2", "stats2", "state2", g["control2"])]:
        res = ol.sirode_unfilter(pars, 1, g["time_start"], g["time"], np.array(g["data"]),
                                 control=ctl)
        assert res["ll"].tolist() == g[ll_key]
        assert res["stats"].tolist() == g[stats_key]
        assert np.array_equal(res["state"][:, 0, :], np.array(g[state_key])[:, -1, :])
        assert np.max(res["stats"][:, 3]) > 1  # steps longer than the reset period
