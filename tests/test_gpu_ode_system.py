"""The continuous-time SIR ODE system on the GPU: against the reference's
own solver driven as a system (tests/golden/sirode_system_dp5.json, written
from dust2/continuous/solver.hpp built in place: states within 1e-10,
identical step counters, step times within 1e-8, critical times and a
re-initialisation after set_state included) and the
reference's system tests (tests/testthat/test-sirode.R)."""
import json
import os

import numpy as np
import pytest
from scipy import stats

import oracle_lib as ol

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "sirode_system_dp5.json")
NAMES = ["N", "I0", "beta", "gamma"]


def golden():
    with open(GOLDEN) as f:
        return json.load(f)["cases"]


def make(lib, pars, P=1, G=1, control=None, critical=None, record=True):
    ctl = lib.dust_ode_control(*(control or [10000, 1e-6, 1e-6, 0, np.inf]),
                               critical_times=critical or None,
                               debug_record_step_times=record)
    plist = [dict(zip(NAMES, p)) for p in np.atleast_2d(pars)]
    sys = lib.dust_system_create(lib.sirode(), plist if G > 1 else plist[0], P, n_groups=G,
                                 ode_control=ctl, deterministic=True)
    lib.dust_system_set_state_initial(sys)
    return sys


@pytest.mark.parametrize("case", golden(), ids=lambda c: c["name"])
def test_matches_reference_solver_as_a_system(gpu_lib, case):
    sys = make(gpu_lib, case["pars"], P=3, control=case["control"],
               critical=case["critical_times"])
    if case["name"] in ("pieces", "restart"):
        out = []
        for i, t in enumerate(case["targets"]):
            gpu_lib.dust_system_run_to_time(sys, t)
            out.append(gpu_lib.dust_system_state(sys))
            if i == case.get("restart_after"):
                # setting the state re-initialises the solver
                gpu_lib.dust_system_set_state(sys, out[-1])
        out = np.array(out)
    else:
        out = gpu_lib.dust_system_simulate(sys, case["targets"])
    want = np.array(case["state"])
    for p in range(3):
        np.testing.assert_allclose(out[:, 0, p, :], want, rtol=1e-10, atol=1e-9)
    for row in gpu_lib.dust_system_internals(sys):
        assert [row["n_steps"], row["n_steps_accepted"], row["n_steps_rejected"]] == \
            case["stats"][:3]
        # the step sizes carry last-place differences of device and host pow()
        assert row["step_size"] == pytest.approx(case["stats"][3], rel=1e-8)
        np.testing.assert_allclose(row["step_times"], case["step_times"], rtol=1e-8)
    assert gpu_lib.dust_system_time(sys) == case["targets"][-1]


def test_can_run_sirode_model(gpu_lib):
    # test-sirode.R:1-15
    sys = make(gpu_lib, [1000, 10, 0.2, 0.1], control=[10000, 1e-6, 1e-6, 0, 0.8])
    res = gpu_lib.dust_system_simulate(sys, np.arange(0.0, 11.0))[:, 0, 0, :]
    assert np.all(np.diff(res[:, 3]) >= 0)
    np.testing.assert_allclose(np.cumsum(res[:, 4]), res[:, 3], rtol=1e-6)


def test_multiple_resets_within_a_step(gpu_lib):
    # test-sirode.R:18-44: running to t in one call matches the daily
    # simulation, and some steps are longer than the reset period
    pars = [1000, 10, 0.2, 0.1]
    ans, long_step = [], False
    for t in range(1, 11):
        sys = make(gpu_lib, pars)
        gpu_lib.dust_system_run_to_time(sys, float(t))
        times = gpu_lib.dust_system_internals(sys)[0]["step_times"]
        long_step |= bool(np.diff(times[-2:])[0] > 1)
        ans.append(gpu_lib.dust_system_state(sys)[0, 0, 4])
    cmp = gpu_lib.dust_system_simulate(make(gpu_lib, pars), np.arange(1.0, 11.0))
    np.testing.assert_allclose(ans, cmp[:, 0, 0, 4], rtol=1e-5)
    assert long_step


def test_run_failure_locks_the_system(gpu_lib):
    # test-sirode.R:67-103
    sys = make(gpu_lib, [1000, 10, 0.2, 0.1], P=10, control=[5, 1e-6, 1e-6, 0, np.inf])
    with pytest.raises(gpu_lib.DustError) as e:
        gpu_lib.dust_system_run_to_time(sys, 30)
    msg = str(e.value)
    assert msg.startswith("10 particles reported errors.")
    assert "1: too many steps" in msg and "- (and 6 more)" in msg
    for fn, args, action in [(gpu_lib.dust_system_run_to_time, (30,), "run"),
                             (gpu_lib.dust_system_time, (), "get time from"),
                             (gpu_lib.dust_system_state, (), "get state from"),
                             (gpu_lib.dust_system_reorder, (np.arange(10),), "reorder"),
                             (gpu_lib.dust_system_set_time, (0,), "set time for"),
                             (gpu_lib.dust_system_simulate, (np.arange(11.0),), "simulate")]:
        with pytest.raises(gpu_lib.DustError,
                           match=f"Can't currently {action} this system: errors are pending"):
            fn(sys, *args)
    gpu_lib.dust_system_update_pars(sys, {"gamma": 0.2})
    gpu_lib.dust_system_set_state_initial(sys)
    assert gpu_lib.dust_system_time(sys) == 0
    with pytest.raises(gpu_lib.DustError,
                       match="Can't run to time -1.000000, system already at time 0.000000"):
        gpu_lib.dust_system_run_to_time(sys, -1)


def test_compare_to_data(gpu_lib):
    # test-sirode.R:105-117
    sys = make(gpu_lib, [1000, 10, 0.1, 0.2], P=10)
    rng = np.random.default_rng(1)
    s = np.zeros((1, 10, 5))
    s[0, :, 4] = rng.poisson(30, 10) + rng.normal(size=10)
    gpu_lib.dust_system_set_state(sys, s)
    got = gpu_lib.dust_system_compare_data(sys, [30.0])
    np.testing.assert_allclose(got[0], stats.poisson.logpmf(30, s[0, :, 4]), rtol=1e-12)


def test_force_solver_to_stop_at_times(gpu_lib):
    # test-sirode.R:119-142
    pars = [1000, 10, 0.2, 0.1]
    s1 = make(gpu_lib, pars, critical=[10.0, 20.0, 30.0])
    s2 = make(gpu_lib, pars)
    gpu_lib.dust_system_run_to_time(s1, 50)
    gpu_lib.dust_system_run_to_time(s2, 50)
    t1 = gpu_lib.dust_system_internals(s1)[0]["step_times"]
    t2 = gpu_lib.dust_system_internals(s2)[0]["step_times"]
    assert all(t in t1 for t in (10, 20, 30))
    assert not all(t in t2 for t in (10, 20, 30))
    np.testing.assert_allclose(gpu_lib.dust_system_state(s1), gpu_lib.dust_system_state(s2),
                               rtol=1e-6)


def test_groups_reorder_and_restarts(gpu_lib):
    rng = np.random.default_rng(3)
    G, P = 4, 6
    pars = np.column_stack([np.full(G, 1000.0), np.full(G, 10.0), rng.uniform(0.1, 0.5, G),
                            rng.uniform(0.05, 0.2, G)])
    # groups of one system equal the unfilter's integration of each set
    time = np.arange(1.0, 31.0)
    sys = make(gpu_lib, pars, P=P, G=G, record=False)
    out = gpu_lib.dust_system_simulate(sys, time)
    ref = ol.sirode_unfilter(pars, 1, 0.0, time, np.zeros(30), n_threads=4)
    for p in range(P):
        np.testing.assert_allclose(out[-1, :, p, :], ref["state"][:, 0, :], rtol=1e-10,
                                   atol=1e-9)
    # perturb particles so that they differ, then reorder halfway: each
    # particle carries its solver state with it
    a = make(gpu_lib, pars, P=P, G=G, record=False)
    b = make(gpu_lib, pars, P=P, G=G, record=False)
    st = gpu_lib.dust_system_state(a)
    st[:, :, 1] += np.arange(P)
    for x in (a, b):
        gpu_lib.dust_system_set_state(x, st)
        gpu_lib.dust_system_run_to_time(x, 5.0)
    perm = np.array([rng.permutation(P) for _ in range(G)])
    gpu_lib.dust_system_reorder(a, perm)
    gpu_lib.dust_system_run_to_time(a, 12.0)
    gpu_lib.dust_system_run_to_time(b, 12.0)
    sb = gpu_lib.dust_system_state(b)
    np.testing.assert_array_equal(gpu_lib.dust_system_state(a),
                                  np.take_along_axis(sb, perm[:, :, None], axis=1))
    # new parameters re-initialise the solver from the current state
    gpu_lib.dust_system_update_pars(b, [{"beta": 0.0}] * G)
    before = gpu_lib.dust_system_state(b)
    gpu_lib.dust_system_run_to_time(b, 20.0)
    after = gpu_lib.dust_system_state(b)
    np.testing.assert_allclose(after[:, :, 0], before[:, :, 0], rtol=1e-12)
    with pytest.raises(ValueError, match="Can't use 'dt' with continuous-time systems"):
        gpu_lib.dust_system_create(gpu_lib.sirode(), dict(zip(NAMES, pars[0])), 1, dt=1)


def test_unfilter_critical_times_match_reference(gpu_lib):
    case = next(c for c in golden() if c["name"] == "critical_times")
    ctl = gpu_lib.dust_ode_control(critical_times=case["critical_times"])
    obj = gpu_lib.dust_unfilter_create(gpu_lib.sirode(), 0.0,
                                       {"time": case["targets"], "incidence": [np.nan]},
                                       ode_control=ctl)
    obj.run(dict(zip(NAMES, case["pars"])))
    np.testing.assert_allclose(obj.last_state()[0, 0], case["state"][-1], rtol=1e-10)
    stats_, _ = obj.last_stats()
    assert list(stats_[0, :3]) == case["stats"][:3]
