"""Pins the golden vectors of the continuous-time models with events,
special variables, resets and delays (tests/golden/ode_general_dp5.json,
written by the reference's own solver, root finder, history and delays
built in place): the closed forms of the reference's tests
(tests/testthat/test-zzz-events.R, test-zzz-special.R), an independent
integration of the SIR ODE and of the delayed logistic equation and, where
the reference tree is built into oracle/_ref, a fresh reference run value
for value."""
import json
import os

import numpy as np
import pytest
from scipy.integrate import solve_ivp

import oracle_lib as ol

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "ode_general_dp5.json")


def cases():
    with open(GOLDEN) as f:
        return {c["name"]: c for c in json.load(f)["cases"]}


def bounce_analytic(t, v0, damp, g=9.8, h0=0.0):
    # tests/testthat/helper-dust.R:91-99, started from height 0
    assert h0 == 0
    t0 = [0.0]
    while t0[-1] < t[-1]:
        t0.append(t0[-1] + 2 * v0 * damp ** (len(t0) - 1) / g)
    t0 = np.array(t0)
    i = np.searchsorted(t0, t, side="right") - 1
    return v0 * damp ** i * (t - t0[i]) - 0.5 * g * (t - t0[i]) ** 2, t0[(t0 > 0) & (t0 < t[-1])]


def test_bounce_roots_and_heights():
    c = cases()["bounce_default"]
    t = np.array(c["targets"])
    y, roots = bounce_analytic(t, 10.0, 0.9)
    ev = np.array(c["events"])
    np.testing.assert_allclose(ev[:, 0], roots, rtol=1e-6)
    np.testing.assert_array_equal(ev[:, 1:], [[0, -1]] * 3)
    np.testing.assert_allclose(np.array(c["state"])[:, 0], y, atol=1e-5)
    # the solver stops at every root: each root is a history step boundary
    h = np.array(c["history"])
    assert all(r in h[:, 0] and r in h[:, 1] for r in ev[:, 0])


def test_change_and_coincident_events():
    for name in ["change_times", "change_coincident"]:
        c = cases()[name]
        tc = np.array(c["pars"]["t_change"])
        dl = np.array(c["pars"]["delta"])
        ev = np.array(c["events"])
        np.testing.assert_allclose(ev[:, 0], tc, rtol=1e-12)
        np.testing.assert_array_equal(ev[:, 1], np.arange(3))
        np.testing.assert_array_equal(ev[:, 2], 1)
        t = np.array(c["targets"])
        want = 0.2 * t + np.concatenate([[0], np.cumsum(dl)])[np.searchsorted(tc, t, "right")]
        np.testing.assert_allclose(np.array(c["state"])[:, 0], want, atol=1e-12)


def test_eventzero_closed_form():
    c = cases()["eventzero_y"]
    np.testing.assert_allclose(c["state"][0], [5000, 99.5], rtol=1e-12)
    np.testing.assert_allclose(np.array(c["events"])[:, 0], [np.sqrt(2 * 6.3456)], rtol=1e-6)
    assert c["stats"][0] < 20
    # y = t^2 / 2; z is reset at the start of the first step after each
    # multiple of 2 (solver.hpp:437-465), so a value reported at a reset
    # time is still the pre-reset one: z = (t^2 - b^2) / 2 with b the last
    # multiple of 2 strictly before t, 1 + t^2 / 2 up to t = 2
    c = cases()["eventzero_both"]
    t = np.array(c["targets"])
    y = np.array(c["state"])
    np.testing.assert_allclose(y[:, 0], t ** 2 / 2, rtol=1e-10, atol=1e-12)
    b = 2 * (np.ceil(t / 2) - 1)
    z = np.where(t <= 2, 1 + t ** 2 / 2, (t ** 2 - b ** 2) / 2)
    np.testing.assert_allclose(y[:, 1], z, rtol=1e-9, atol=1e-9)
    # value_z = 1.5 is crossed upward once before the first reset (t = 1)
    # and then within every later period in which z reaches it
    ev = np.array(c["events"])
    assert ev[0, 1] == 1 and ev[0, 0] == pytest.approx(1.0, rel=1e-6)
    assert 0 in ev[:, 1]  # y crosses 30 at t = sqrt(60)
    assert ev[ev[:, 1] == 0, 0] == pytest.approx(np.sqrt(60), rel=1e-6)


def test_specialevent_closed_form():
    t = np.linspace(0, 10, 101)
    y0 = np.array(cases()["specialevent_none"]["state"])
    np.testing.assert_allclose(y0[:, 0], t, atol=1e-12)
    np.testing.assert_array_equal(y0[:, 1], 0)
    y3 = np.array(cases()["specialevent_at_3"]["state"])
    np.testing.assert_allclose(y3[:, 0], np.where(t < 3, t, 6 - t), atol=1e-10)
    np.testing.assert_array_equal(y3[:, 1], (t >= 3).astype(float))


def test_sirdelay_against_independent_integration():
    for name in ["sirdelay_default", "sirdelay_fast"]:
        c = cases()[name]
        p = dict({"N": 1000.0, "I0": 10.0, "beta": 0.2, "gamma": 0.1}, **c["pars"])

        def f(_, y):
            si = p["beta"] * y[0] * y[1] / p["N"]
            return [-si, si - p["gamma"] * y[1], p["gamma"] * y[1], si]

        t = np.array(c["targets"])
        grid = np.union1d(t, np.maximum(t - 1, 0))
        sol = solve_ivp(f, (0, t[-1]), [p["N"] - p["I0"], p["I0"], 0, 0], t_eval=grid,
                        rtol=1e-11, atol=1e-11, method="DOP853")
        cum = dict(zip(grid, sol.y[3]))
        y = np.array(c["state"])
        np.testing.assert_allclose(y[:, :4], sol.y[:, np.searchsorted(grid, t)].T, rtol=2e-5,
                                   atol=1e-5)
        inc = np.array([cum[x] - cum[max(x - 1, 0)] for x in t])
        np.testing.assert_allclose(y[:, 4], inc, rtol=2e-5, atol=1e-5)


def test_logdelay_against_method_of_steps():
    for name in ["logdelay_oscillating", "logdelay_tight"]:
        c = cases()[name]
        p = dict({"r": 1.0, "K": 100.0, "tau": 1.0, "y0": 10.0}, **c["pars"])
        # classical RK4 on a grid that divides tau, the delayed value read
        # from the grid itself
        n = 400
        dt = p["tau"] / n
        t_end = c["targets"][-1]
        m = int(round(t_end / dt))
        y = np.empty(m + 1)
        y[0] = p["y0"]

        def lag(k2):  # y at grid index k2 (half steps by averaging)
            if k2 < 0:
                return p["y0"]
            i = int(np.floor(k2))
            return y[i] if k2 == i else 0.5 * (y[i] + y[i + 1])

        for k in range(m):
            d0, dh, d1 = lag(k - n), lag(k + 0.5 - n), lag(k + 1 - n)
            g = lambda yy, d: p["r"] * yy * (1 - d / p["K"])  # noqa: E731
            k1 = g(y[k], d0)
            k2 = g(y[k] + dt / 2 * k1, dh)
            k3 = g(y[k] + dt / 2 * k2, dh)
            k4 = g(y[k] + dt * k3, d1)
            y[k + 1] = y[k] + dt / 6 * (k1 + 2 * k2 + 2 * k3 + k4)
        idx = np.round(np.array(c["targets"]) / dt).astype(int)
        np.testing.assert_allclose(np.array(c["state"])[:, 0], y[idx], rtol=2e-4, atol=1e-4)
        # before tau the delayed value is the initial one: pure exponential
        t = np.array(c["targets"])
        early = t <= p["tau"]
        np.testing.assert_allclose(
            np.array(c["state"])[early, 0],
            p["y0"] * np.exp(p["r"] * (1 - p["y0"] / p["K"]) * t[early]), rtol=1e-5)
        # no step is longer than the delay (delays.hpp:85-92)
        assert np.max(np.diff(c["step_times"])) <= p["tau"] + 1e-12


@pytest.mark.skipif(not ol.ref_available(), reason="reference tree not built into oracle/_ref")
@pytest.mark.parametrize("name", sorted(cases()))
def test_golden_is_the_reference_run(name):
    c = cases()[name]
    r = ol.ref_ode_general(c["model"], c["row"], c["targets"], control=c["control"],
                           critical_times=c["critical_times"], save_history=c["save_history"])
    np.testing.assert_array_equal(r["state"], c["state"])
    np.testing.assert_array_equal(r["stats"], c["stats"])
    np.testing.assert_array_equal(r["events"].reshape(-1), np.array(c["events"]).reshape(-1))
