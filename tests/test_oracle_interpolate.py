"""Pins the oracle's interpolation (piecewise constant, piecewise linear,
natural cubic spline): the reference's own test cases
(tests/testthat/test-interpolate.R), closed forms (numpy's interp, a
direct lookup, scipy's natural cubic spline) and, where the reference
tree is present, the reference's interpolators compiled in place
(oracle/_ref), value for value."""
import numpy as np
import pytest

import oracle_lib as ol

KINDS = ["constant", "linear", "spline"]


def test_search_positions_and_errors():
    # test-interpolate.R:1-29: the interval index, read through piecewise
    # constant interpolation of y = index
    t = np.arange(10) + 0.5
    eps = 1e-7
    idx = np.arange(10.0)
    cases = [(0.5, 0), (0.5 + eps, 0), (1.5 - eps, 0), (1.5, 1), (1.5 + eps, 1),
             (2.5 - eps, 1), (5.5, 5), (5.5 + eps, 5), (6.5 - eps, 5), (9.5 - eps, 8),
             (9.5, 9)]
    for z, i in cases:
        assert ol.interpolate("constant", t, idx, [z])[0] == i
    with pytest.raises(ol.OracleError, match="before"):
        ol.interpolate("constant", t, idx, [0.0])
    # beyond the last time: constant extrapolates, the others refuse
    assert ol.interpolate("constant", t, idx, [9.5 + eps])[0] == 9
    for kind in ["linear", "spline"]:
        with pytest.raises(ol.OracleError, match="after"):
            ol.interpolate(kind, t, idx, [9.5 + eps])
    assert ol.interpolate("linear", t, idx, [9.5])[0] == 9


def test_constant_linear_closed_forms():
    # test-interpolate.R:32-74
    rng = np.random.default_rng(1)
    t = np.arange(11.0)
    y = rng.random(11)
    z = np.concatenate([t, rng.uniform(0, 10, 200), [10 - 1e-8, 1 - 1e-8]])
    const = ol.interpolate("constant", t, y, z)
    np.testing.assert_array_equal(const, y[np.searchsorted(t, z, side="right") - 1])
    np.testing.assert_array_equal(ol.interpolate("constant", t, y, [100.0]), y[-1:])
    lin = ol.interpolate("linear", t, y, z)
    np.testing.assert_allclose(lin, np.interp(z, t, y), rtol=1e-14, atol=1e-15)
    np.testing.assert_array_equal(ol.interpolate("linear", t, y, t), y)


def test_spline_matches_natural_cubic_spline():
    # test-interpolate.R:76-98: R's spline(method = "natural")
    from scipy.interpolate import CubicSpline
    rng = np.random.default_rng(2)
    for n in [2, 3, 5, 20]:
        t = np.cumsum(rng.uniform(0.3, 2.0, n))
        y = rng.normal(size=n)
        z = np.concatenate([t, rng.uniform(t[0], t[-1], 100)])
        got = ol.interpolate("spline", t, y, z)
        want = CubicSpline(t, y, bc_type="natural")(z)
        np.testing.assert_allclose(got, want, rtol=1e-11, atol=1e-12)
    # a quadratic is not reproduced (natural ends), a straight line is
    t = np.arange(6.0)
    np.testing.assert_allclose(ol.interpolate("spline", t, 2 * t + 1, [0.3, 2.7, 4.9]),
                               [1.6, 6.4, 10.8], rtol=1e-14)


def test_matrix_series_match_single_series():
    # test-interpolate.R:112-160: every series of an array interpolator
    # equals the single-series interpolator of that series
    rng = np.random.default_rng(3)
    t = np.cumsum(rng.uniform(0.5, 1.5, 8))
    y = rng.normal(size=(8, 3))
    z = rng.uniform(t[0], t[-1], 50)
    for kind in KINDS:
        arr = ol.interpolate(kind, t, y, z)
        for j in range(3):
            one = ol.interpolate(kind, t, y[:, j], z)
            np.testing.assert_allclose(arr[:, j], one, rtol=1e-14, atol=1e-15)


@pytest.mark.skipif(not ol.ref_available(), reason="reference tree not built into oracle/_ref")
@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("n_series", [1, 3])
def test_oracle_equals_reference_build(kind, n_series):
    rng = np.random.default_rng(4 + n_series)
    t = np.cumsum(rng.uniform(0.1, 3.0, 30))
    y = rng.normal(size=(30, n_series)) if n_series > 1 else rng.normal(size=30)
    z = np.concatenate([t, rng.uniform(t[0], t[-1], 300)])
    np.testing.assert_array_equal(ol.interpolate(kind, t, y, z),
                                  ol.ref_interpolate(kind, t, y, z))


# ---------------------------------------------------------------------------
# The SIR with beta interpolated in time (oracle model sir_tv)
SIR = [1000.0, 10.0, 0.0, 0.1, 1e6]


@pytest.mark.parametrize("kind", KINDS)
def test_sir_tv_flat_schedule_is_the_plain_sir(kind):
    # a schedule that is 0.3 everywhere reproduces the SIR with beta = 0.3
    # draw for draw, whatever the interpolation
    time = np.arange(1.0, 31.0)
    data = np.arange(30.0)[:, None]
    beta_t = np.array([0.0, 7.0, 15.0, 40.0])
    rows = ol.sir_tv_rows([SIR], kind, beta_t, np.full(4, 0.3))
    plain = np.array([[1000.0, 10.0, 0.3, 0.1, 1e6]])
    a = ol.sir_filter(rows, 50, 0.0, time, 0.25, data, seed=3, model="sir_tv")
    b = ol.sir_filter(plain, 50, 0.0, time, 0.25, data, seed=3)
    np.testing.assert_array_equal(a["ll"], b["ll"])
    np.testing.assert_array_equal(a["state"], b["state"])


def test_sir_tv_uses_beta_at_step_start_time():
    # beta drops to 0 from time 10: no infection in a step starting at or
    # after 10, so S at time 30 equals S at time 10 in every particle
    rows = ol.sir_tv_rows([[1000.0, 10.0, 0.0, 0.05, 1e6]], "constant",
                          np.array([0.0, 10.0]), np.array([0.8, 0.0]))
    out, _ = ol.sir_simulate(rows, 40, 0.0, 0.25, np.array([10.0, 30.0]),
                             ol.prng(40, seed=5), model="sir_tv")
    np.testing.assert_array_equal(out[1, 0, :, 0], out[0, 0, :, 0])
    assert np.all(out[0, 0, :, 0] < 990)


def test_sir_tv_reports_times_outside_the_knots():
    rows = ol.sir_tv_rows([SIR], "linear", np.array([0.0, 9.0]), np.array([0.2, 0.3]))
    with pytest.raises(ol.OracleError,
                       match=r"Tried to interpolate at time = 9.25, which is 0.25 after the "
                             r"last time \(9\)"):
        ol.sir_filter(rows, 10, 0.0, np.arange(1.0, 11.0), 0.25, np.ones((10, 1)), seed=1,
                      model="sir_tv")
