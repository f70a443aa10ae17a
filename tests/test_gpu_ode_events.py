"""Continuous-time systems with events, carried (special) variables,
periodic resets and delays on the GPU: against the reference's own solver,
root finder, history and delays built in place
(tests/golden/ode_general_dp5.json: states within 1e-10, identical step
counters and events, history within 1e-10, step sizes within 1e-8), and the
reference's tests of these features (tests/testthat/test-zzz-events.R,
test-zzz-special.R, test-zzz-compile.R:101-117)."""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "ode_general_dp5.json")


def golden():
    with open(GOLDEN) as f:
        return json.load(f)["cases"]


def make(lib, model, pars, P=1, G=1, control=None, critical=(), save_history=False,
         record=True, **kwargs):
    ctl = lib.dust_ode_control(*(control or [10000, 1e-6, 1e-6, 0, np.inf]),
                               critical_times=list(critical) or None,
                               debug_record_step_times=record, save_history=save_history)
    gen = getattr(lib, model)()
    sys = lib.dust_system_create(gen, pars, P, n_groups=G, ode_control=ctl, **kwargs)
    lib.dust_system_set_state_initial(sys)
    return sys


def check_particle(row, case):
    assert [row["n_steps"], row["n_steps_accepted"], row["n_steps_rejected"]] == \
        case["stats"][:3]
    # step sizes and times carry last-place differences of device and host
    # pow() in the controller, amplified by the cancellation in the error
    # estimate at tight tolerances; states stay within 1e-10
    assert row["step_size"] == pytest.approx(case["stats"][3], rel=1e-8)
    np.testing.assert_allclose(row["step_times"], case["step_times"], rtol=1e-8)
    ev = np.array(case["events"]).reshape(-1, 3)
    if "events" in row:
        np.testing.assert_allclose(row["events"]["time"], ev[:, 0], rtol=1e-10)
        np.testing.assert_array_equal(row["events"]["index"], ev[:, 1].astype(int) + 1)
        np.testing.assert_array_equal(row["events"]["sign"], ev[:, 2])
    else:
        assert ev.shape[0] == 0
    if case["history"] is not None:
        h = np.array(case["history"])
        got = row["history"]
        assert got["n_saved"] == h.shape[0]
        np.testing.assert_allclose(got["t0"], h[:, 0], rtol=1e-10)
        np.testing.assert_allclose(got["t1"], h[:, 1], rtol=1e-10)
        np.testing.assert_allclose(got["h"], h[:, 2], rtol=1e-10)
        H = got["coefficients"].shape[0]
        coef = h[:, 3:].reshape(-1, 5, H).transpose(2, 1, 0)
        np.testing.assert_allclose(got["coefficients"], coef, rtol=1e-9, atol=1e-12)


@pytest.mark.parametrize("case", golden(), ids=lambda c: c["name"])
def test_matches_reference_solver(gpu_lib, case):
    sys = make(gpu_lib, case["model"], case["pars"], P=2, control=case["control"],
               critical=case["critical_times"], save_history=case["save_history"])
    out = gpu_lib.dust_system_simulate(sys, case["targets"])
    want = np.array(case["state"])
    for p in range(2):
        np.testing.assert_allclose(out[:, 0, p, :], want, rtol=1e-10, atol=1e-10)
    for row in gpu_lib.dust_system_internals(sys, include_history=True):
        check_particle(row, case)


def test_groups_match_their_own_runs(gpu_lib):
    # bounce parameter sets as groups of one system: each group equals a
    # system of that set alone, and its particles are identical
    sets = [{}, {"damp": 0.7, "velocity": 5.0, "height": 2.0}, {"damp": 0.95}]
    t = np.linspace(0, 6, 25)
    sys = make(gpu_lib, "bounce", sets, P=3, G=len(sets), save_history=True)
    out = gpu_lib.dust_system_simulate(sys, t)
    rows = gpu_lib.dust_system_internals(sys, include_history=True)
    for g, p in enumerate(sets):
        one = make(gpu_lib, "bounce", p, save_history=True)
        np.testing.assert_array_equal(out[:, g], np.repeat(
            gpu_lib.dust_system_simulate(one, t)[:, 0], 3, axis=1))
        ev = gpu_lib.dust_system_internals(one)[0]["events"]
        for k in range(3):
            np.testing.assert_array_equal(rows[3 * g + k]["events"]["time"], ev["time"])


def r_all_equal(current, target, tolerance):
    # testthat's expect_equal(tolerance =) through all.equal.numeric: the
    # mean absolute difference relative to the mean absolute target
    current, target = np.asarray(current), np.asarray(target)
    xy = np.mean(np.abs(target - current))
    xn = np.mean(np.abs(target))
    return (xy / xn if np.isfinite(xn) and xn > tolerance else xy) < tolerance


def example_bounce_analytic(t, v0=10.0, damp=0.9, g=9.8):
    # tests/testthat/helper-dust.R:91-99
    t0 = [0.0]
    while t0[-1] < t[-1]:
        t0.append(t0[-1] + 2 * v0 * damp ** (len(t0) - 1) / g)
    t0 = np.array(t0)
    i = np.searchsorted(t0, t, side="right") - 1
    y = v0 * damp ** i * (t - t0[i]) - 0.5 * g * (t - t0[i]) ** 2
    return y, t0[(t0 > 0) & (t0 < t[-1])]


def test_roots_and_events(gpu_lib):
    # test-zzz-events.R:1-30
    sys = make(gpu_lib, "bounce", {}, save_history=True)
    t = np.linspace(0, 6, 60)
    y = gpu_lib.dust_system_simulate(sys, t)
    y_cmp, roots = example_bounce_analytic(t)
    info = gpu_lib.dust_system_internals(sys, include_history=True)[0]
    r = info["events"]
    assert len(r["time"]) == 3
    assert r["name"] == ["bounce"] * 3
    np.testing.assert_allclose(r["time"], roots, rtol=1e-6)
    np.testing.assert_array_equal(r["index"], [1, 1, 1])
    np.testing.assert_array_equal(r["sign"], [-1, -1, -1])
    h = info["history"]
    assert all(x in h["t0"] for x in r["time"])
    assert all(x in h["t1"] for x in r["time"])
    assert r_all_equal(y[:, 0, 0, 0], y_cmp, 1e-6)
    np.testing.assert_allclose(y[:, 0, 0, 0], y_cmp, atol=1e-5)


@pytest.mark.parametrize("t_change,delta", [([2.0, 5.1234, 7.0], [0.4, -1.1, 2.3]),
                                            ([2.0, 2.0, 3.0], [1.0, 3.0, 5.0])])
def test_events_in_time(gpu_lib, t_change, delta):
    # test-zzz-events.R:33-79, the second a pair of coincident events
    sys = make(gpu_lib, "change", {"r": 0.2, "n": 3, "t_change": t_change, "delta": delta},
               save_history=True)
    t = np.linspace(0, 10, 101)
    y = gpu_lib.dust_system_simulate(sys, t)
    ev = gpu_lib.dust_system_internals(sys)[0]["events"]
    assert ev["name"] == ["event-1", "event-2", "event-3"]
    np.testing.assert_allclose(ev["time"], t_change, rtol=1e-12)
    np.testing.assert_array_equal(ev["index"], [1, 2, 3])
    np.testing.assert_array_equal(ev["sign"], [1, 1, 1])
    expected = t * 0.2 + np.concatenate([[0], np.cumsum(delta)])[
        np.searchsorted(t_change, t, side="right")]
    np.testing.assert_allclose(y[:, 0, 0, 0], expected, rtol=1e-10, atol=1e-12)


def test_events_that_do_not_alter_the_root(gpu_lib):
    # test-zzz-events.R:82-96: the solution is t^2 / 2r
    sys = make(gpu_lib, "eventzero", {"value_y": 6.3456}, save_history=True)
    gpu_lib.dust_system_run_to_time(sys, 100)
    assert gpu_lib.dust_system_time(sys) == 100
    np.testing.assert_allclose(gpu_lib.dust_system_state(sys)[0, 0], [5000, 99.5], rtol=1e-12)
    info = gpu_lib.dust_system_internals(sys, include_history=True)[0]
    np.testing.assert_allclose(info["events"]["time"], [np.sqrt(2 * 6.3456)], rtol=1e-6)
    assert info["n_steps"] < 20


def test_special_variables_and_events(gpu_lib):
    # test-zzz-special.R:26-42
    t = np.linspace(0, 10, 101)
    sys = make(gpu_lib, "specialevent", {})
    y = gpu_lib.dust_system_simulate(sys, t)[:, 0, 0]
    np.testing.assert_allclose(y[:, 0], t, rtol=1e-12, atol=1e-12)
    np.testing.assert_array_equal(y[:, 1], 0)
    sys = make(gpu_lib, "specialevent", {"time_activate": 3})
    y = gpu_lib.dust_system_simulate(sys, t)[:, 0, 0]
    np.testing.assert_allclose(y[:, 0], np.where(t < 3, t, 6 - t), rtol=1e-10, atol=1e-10)
    np.testing.assert_array_equal(y[:, 1], (t >= 3).astype(float))


def test_delay_model_equals_the_sir_ode(gpu_lib):
    # test-zzz-compile.R:101-117: the SIR with a delayed incidence equals the
    # SIR ODE with the incidence reset daily
    t = np.arange(0.0, 101.0, 4.0)
    cmp = gpu_lib.dust_system_create(gpu_lib.sirode(), {}, 1)
    gpu_lib.dust_system_set_state_initial(cmp)
    y_cmp = gpu_lib.dust_system_simulate(cmp, t)[:, 0, 0]
    sys = make(gpu_lib, "sirdelay", {})
    y = gpu_lib.dust_system_simulate(sys, t)[:, 0, 0]
    np.testing.assert_allclose(y[:, :4], y_cmp[:, :4], rtol=1e-5)
    np.testing.assert_allclose(y[:, 4], y_cmp[:, 4], rtol=1e-5, atol=1e-8)


def test_delayed_rhs_history_and_its_limits(gpu_lib):
    # the delayed logistic: before tau the delayed value is the initial one,
    # so y grows as y0 exp(r (1 - y0/K) t) on [0, tau]
    pars = {"r": 1.8, "K": 100.0, "tau": 1.0, "y0": 10.0}
    sys = make(gpu_lib, "logdelay", pars, record=False)
    y = gpu_lib.dust_system_simulate(sys, [0.5, 1.0])[:, 0, 0, 0]
    np.testing.assert_allclose(y, 10 * np.exp(1.8 * 0.9 * np.array([0.5, 1.0])), rtol=1e-5)
    # no step is longer than the delay (delays.hpp:85-92)
    sys = make(gpu_lib, "logdelay", pars)
    gpu_lib.dust_system_run_to_time(sys, 30.0)
    st = gpu_lib.dust_system_internals(sys)[0]["step_times"]
    assert np.max(np.diff(st)) <= 1.0 + 1e-12
    # a history ring too short for the delay is reported, and locks the system
    short = make(gpu_lib, "logdelay", pars, history_cap=2)
    with pytest.raises(gpu_lib.DustError, match="History too short for a delayed value"):
        gpu_lib.dust_system_run_to_time(short, 10.0)
    with pytest.raises(gpu_lib.DustError, match="errors are pending"):
        gpu_lib.dust_system_state(short)


def test_reorder_update_pars_and_errors(gpu_lib):
    P = 4
    sys = make(gpu_lib, "bounce", {}, P=P, save_history=True)
    ref = make(gpu_lib, "bounce", {}, P=P, save_history=True)
    st = gpu_lib.dust_system_state(sys)
    st[0, :, 1] = [10.0, 8.0, 6.0, 12.0]  # different launch velocities
    for s in (sys, ref):
        gpu_lib.dust_system_set_state(s, st)
        gpu_lib.dust_system_run_to_time(s, 3.0)
    perm = np.array([[2, 0, 3, 1]])
    gpu_lib.dust_system_reorder(sys, perm)
    for s in (sys, ref):
        gpu_lib.dust_system_run_to_time(s, 7.0)
    np.testing.assert_array_equal(gpu_lib.dust_system_state(sys)[0],
                                  gpu_lib.dust_system_state(ref)[0][perm[0]])
    a = gpu_lib.dust_system_internals(sys, include_history=True)
    b = gpu_lib.dust_system_internals(ref, include_history=True)
    for k in range(P):
        j = perm[0, k]
        np.testing.assert_array_equal(a[k]["events"]["time"], b[j]["events"]["time"])
        np.testing.assert_array_equal(a[k]["history"]["t1"], b[j]["history"]["t1"])
        np.testing.assert_array_equal(a[k]["step_times"], b[j]["step_times"])
    # a new damping re-initialises the solver: the run equals a fresh system
    # started from the same state and time with that damping
    gpu_lib.dust_system_update_pars(sys, {"damp": 0.5})
    fresh = make(gpu_lib, "bounce", {"damp": 0.5}, P=P, save_history=True)
    gpu_lib.dust_system_set_state(fresh, gpu_lib.dust_system_state(sys))
    gpu_lib.dust_system_set_time(fresh, 7.0)
    t = np.linspace(7.0, 9.0, 21)
    np.testing.assert_array_equal(gpu_lib.dust_system_simulate(sys, t),
                                  gpu_lib.dust_system_simulate(fresh, t))
    assert gpu_lib.dust_system_internals(sys)[0]["step_times"][0] == 7.0
    # too many steps: every particle reports, the system locks
    bad = make(gpu_lib, "bounce", {}, P=3, control=[20, 1e-6, 1e-6, 0, 0.01])
    with pytest.raises(gpu_lib.DustError) as e:
        gpu_lib.dust_system_run_to_time(bad, 5.0)
    assert str(e.value).startswith("3 particles reported errors.")
    assert "too many steps" in str(e.value)
    with pytest.raises(gpu_lib.DustError, match="Can't currently run this system"):
        gpu_lib.dust_system_run_to_time(bad, 6.0)
    with pytest.raises(ValueError, match="At most 16 changes"):
        make(gpu_lib, "change", {"r": 1, "t_change": np.arange(17.0), "delta": np.ones(17)})
