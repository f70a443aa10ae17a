"""GPU parity of the Dormand-Prince 5(4) unfilter (SIR ODE).

FP64: log-likelihoods within 1e-10 relative of the reference solver's golden
vectors (tests/golden/sirode_dp5.json, written from the reference's own
dust2/continuous/solver.hpp), identical step counters, states within 1e-10.
FP32 state with FP64 error norm and controller: states and log-likelihoods
within the solver tolerance of the FP64 oracle.
"""
import json
import os

import numpy as np
import pytest

import oracle_lib as ol

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "sirode_dp5.json")


def golden():
    with open(GOLDEN) as f:
        return json.load(f)


def make_unfilter(lib, pars, time, data, P=1, control=None, precision="fp64"):
    G = len(pars)
    obj = lib.dust_unfilter_create(lib.sirode(), 0.0, {"time": time, "incidence": data},
                                   n_particles=P, n_groups=G, ode_control=control,
                                   precision=precision)
    plist = [dict(zip(["N", "I0", "beta", "gamma"], p)) for p in pars]
    ll = obj.run(plist)
    return obj, np.atleast_1d(ll)


def test_matches_reference_solver_golden(gpu_lib):
    g = golden()
    pars = np.array(g["pars"])
    data = np.array(g["data"])
    obj, ll = make_unfilter(gpu_lib, pars, g["time"], data)
    np.testing.assert_allclose(ll, g["ll"], rtol=1e-10)
    stats, _ = obj.last_stats()
    np.testing.assert_array_equal(stats[:, :3], np.array(g["stats"])[:, :3])
    np.testing.assert_allclose(stats[:, 3], np.array(g["stats"])[:, 3], rtol=1e-10)
    final = np.array(g["state"])[:, -1, :]
    np.testing.assert_allclose(obj.last_state()[:, 0, :], final, rtol=1e-10, atol=1e-10)


def test_tight_control_matches_reference_solver(gpu_lib):
    g = golden()
    pars = np.array(g["pars"])[:4]
    data = np.array(g["data"][:60])
    obj, ll = make_unfilter(gpu_lib, pars, g["time"][:60], data,
                            control=gpu_lib.dust_ode_control(*g["control2"]))
    np.testing.assert_allclose(ll, g["ll2"], rtol=1e-10)
    stats, _ = obj.last_stats()
    np.testing.assert_array_equal(stats[:, :3], np.array(g["stats2"])[:, :3])


def test_many_parameter_sets_match_oracle(gpu_lib):
    rng = np.random.default_rng(3)
    n = 4096
    pars = np.column_stack([np.full(n, 1000.0), np.full(n, 10.0), rng.uniform(0.1, 0.5, n),
                            rng.uniform(0.05, 0.2, n)])
    time = np.arange(1.0, 366.0)
    data = rng.poisson(6, 365).astype(float)
    obj, ll = make_unfilter(gpu_lib, pars, time, data)
    ref = ol.sirode_unfilter(pars, 1, 0.0, time, data, n_threads=8)
    np.testing.assert_allclose(ll, ref["ll"], rtol=1e-10)
    np.testing.assert_allclose(obj.last_state(), ref["state"], rtol=1e-10, atol=1e-9)


def test_fp32_state_within_tolerance(gpu_lib):
    rng = np.random.default_rng(4)
    n = 2048
    pars = np.column_stack([np.full(n, 1000.0), np.full(n, 10.0), rng.uniform(0.1, 0.5, n),
                            rng.uniform(0.05, 0.2, n)])
    time = np.arange(1.0, 366.0)
    data = rng.poisson(6, 365).astype(float)
    obj, ll = make_unfilter(gpu_lib, pars, time, data, precision="fp32")
    ref = ol.sirode_unfilter(pars, 1, 0.0, time, data, n_threads=8)
    # FP32 state and stages, FP64 error norm and controller, 365 days at
    # rtol 1e-6 per step: S, R and cumulative cases within 1e-5 relative
    # (measured 2.7e-6), I and the daily incidence, which fall to ~1e-3 as
    # epidemics end, within 1e-5 of N absolute (measured 3.6e-6), the
    # log-likelihood within 1e-5 relative (measured 2.5e-6)
    st = obj.last_state()
    for c in [0, 2, 3]:
        np.testing.assert_allclose(st[:, 0, c], ref["state"][:, 0, c], rtol=1e-5)
    for c in [1, 4]:
        np.testing.assert_allclose(st[:, 0, c], ref["state"][:, 0, c], rtol=0, atol=1e-5 * 1000)
    np.testing.assert_allclose(ll, ref["ll"], rtol=1e-5)


def test_replicated_particles_identical(gpu_lib):
    # test-unfilter.R:148-175
    pars = np.array([[1000, 10, 0.3, 0.1], [1000, 10, 0.2, 0.1]])
    time = np.arange(1.0, 31.0)
    data = np.arange(30.0)
    obj, ll = make_unfilter(gpu_lib, pars, time, data, P=4)
    ll = ll.reshape(2, 4)
    assert np.all(ll == ll[:, :1])


def test_too_many_steps_error(gpu_lib):
    # test-sirode.R:67-103
    pars = np.array([[1000, 10, 0.3, 0.1]])
    with pytest.raises(gpu_lib.DustError, match="too many steps"):
        make_unfilter(gpu_lib, pars, np.arange(1.0, 11.0), np.ones(10),
                      control=gpu_lib.dust_ode_control(max_steps=5))
    # every failing particle is counted, on each of two failing runs (the
    # error flag and records are cleared between runs), and only the
    # selected groups' particles when a subset runs
    pars = np.array([[1000, 10, 0.3, 0.1]] * 3)
    time = np.arange(1.0, 11.0)
    obj = gpu_lib.dust_unfilter_create(gpu_lib.sirode(), 0.0,
                                       {"time": time, "incidence": np.ones(10)},
                                       n_particles=2, n_groups=3,
                                       ode_control=gpu_lib.dust_ode_control(max_steps=5))
    plist = [dict(zip(["N", "I0", "beta", "gamma"], p)) for p in pars]
    for _ in range(2):
        with pytest.raises(gpu_lib.DustError) as e:
            obj.run(plist)
        assert str(e.value).startswith("6 particles reported errors.")
        assert "too many steps" in str(e.value)
    with pytest.raises(gpu_lib.DustError, match="^2 particles reported errors."):
        obj.run(None, index_group=[1])


def test_gradient_matches_oracle_and_keeps_likelihood(gpu_lib):
    rng = np.random.default_rng(21)
    G = 256
    pars = np.column_stack([np.full(G, 1000.0), rng.uniform(5, 20, G),
                            rng.uniform(0.1, 0.5, G), rng.uniform(0.05, 0.2, G)])
    time = np.arange(1.0, 121.0)
    data = rng.poisson(8, (120, G)).astype(float)
    data[[7, 50], :] = np.nan
    plist = [dict(zip(["N", "I0", "beta", "gamma"], p)) for p in pars]
    obj = gpu_lib.dust_unfilter_create(gpu_lib.sirode(), 0.0, {"time": time, "incidence": data},
                                       n_particles=1, n_groups=G)
    ll_plain = obj.run(plist)
    with pytest.raises(gpu_lib.DustError, match="System was not run with 'adjoint = TRUE'"):
        obj.last_gradient()
    ll = gpu_lib.dust_likelihood_run(obj, None, adjoint=True)
    grad = gpu_lib.dust_likelihood_last_gradient(obj)
    assert np.array_equal(ll, ll_plain)       # same steps: the plain likelihood
    ref_ll, ref_grad = ol.sirode_gradient(pars, 0.0, time, data, n_threads=8)
    np.testing.assert_allclose(ll, ref_ll, rtol=1e-10)
    np.testing.assert_allclose(grad, ref_grad, rtol=1e-10, atol=1e-10 * np.abs(ref_grad).max())


def test_plain_and_gradient_runs_take_identical_steps(gpu_lib):
    # The gradient run integrates the sensitivities beside the state with
    # the same step sequence as the plain run: likelihoods, final states and
    # trajectories agree bit for bit, with critical times, the daily reset,
    # missing data, a group subset and an odd set count.
    rng = np.random.default_rng(23)
    G = 301
    pars = np.column_stack([rng.uniform(500, 5000, G), rng.uniform(1, 20, G),
                            rng.uniform(0.1, 0.8, G), rng.uniform(0.05, 0.3, G)])
    time = np.concatenate([np.arange(0.5, 30.0, 1.5), np.arange(31.0, 151.0)])
    data = rng.poisson(10, (time.size, G)).astype(float)
    data[[3, 40, 41], :] = np.nan
    plist = [dict(zip(["N", "I0", "beta", "gamma"], p)) for p in pars]
    control = gpu_lib.dust_ode_control(critical_times=[10.25, 60.0, 97.5])
    obj = gpu_lib.dust_unfilter_create(gpu_lib.sirode(), 0.0, {"time": time, "incidence": data},
                                       n_particles=1, n_groups=G, ode_control=control)
    for sel in (None, np.arange(7, G, 3)):
        ll_plain = obj.run(plist if sel is None else None, index_group=sel,
                           save_trajectories=True)
        st_plain = obj.last_state()
        tr_plain = obj.last_trajectories()
        ll_grad = obj.run(None, index_group=sel, adjoint=True, save_trajectories=True)
        assert np.array_equal(ll_plain, ll_grad)
        assert np.array_equal(obj.last_state(), st_plain)
        assert np.array_equal(obj.last_trajectories(), tr_plain)
        assert np.all(np.isfinite(ll_plain))


def test_gradient_matches_central_differences(gpu_lib):
    rng = np.random.default_rng(22)
    G = 32
    pars = np.column_stack([np.full(G, 1000.0), rng.uniform(5, 20, G),
                            rng.uniform(0.1, 0.5, G), rng.uniform(0.05, 0.2, G)])
    time = np.arange(1.0, 61.0)
    data = rng.poisson(8, 60).astype(float)
    obj = gpu_lib.dust_unfilter_create(gpu_lib.sirode(), 0.0, {"time": time, "incidence": data},
                                       n_particles=1, n_groups=G)
    names = ["N", "I0", "beta", "gamma"]
    gpu_lib.dust_likelihood_run(obj, [dict(zip(names, p)) for p in pars], adjoint=True)
    grad = obj.last_gradient()
    fd = np.zeros_like(grad)
    for j, col in enumerate([2, 3, 1]):
        h = 1e-6 * pars[:, col]
        up, dn = pars.copy(), pars.copy()
        up[:, col] += h
        dn[:, col] -= h
        lu = obj.run([dict(zip(names, p)) for p in up])
        ld = obj.run([dict(zip(names, p)) for p in dn])
        fd[:, j] = (lu - ld) / (2 * h)
    np.testing.assert_allclose(grad, fd, rtol=1e-6)


@pytest.mark.parametrize("which", ["default", "loose"])
def test_weekly_data_matches_reference_solver(gpu_lib, which):
    # steps spanning several daily resets (reset by the dense output at the
    # boundary, final correction after the last step), reference goldens
    with open(os.path.join(os.path.dirname(__file__), "golden", "sirode_weekly_dp5.json")) as f:
        g = json.load(f)
    pars = np.array(g["pars"])
    control = None if which == "default" else gpu_lib.dust_ode_control(*g["control2"])
    key = "" if which == "default" else "2"
    obj, ll = make_unfilter(gpu_lib, pars, g["time"], np.array(g["data"]), control=control)
    np.testing.assert_allclose(ll, g["ll" + key], rtol=1e-10)
    stats, _ = obj.last_stats()
    np.testing.assert_array_equal(stats[:, :3], np.array(g["stats" + key])[:, :3])
    np.testing.assert_allclose(obj.last_state()[:, 0, :], np.array(g["state" + key])[:, -1, :],
                               rtol=1e-10, atol=1e-9)


def test_ode_unfilter_trajectories(gpu_lib):
    g = golden()
    pars = np.array(g["pars"])[:3]
    data = np.array(g["data"])
    obj = gpu_lib.dust_unfilter_create(gpu_lib.sirode(), 0.0,
                                       {"time": g["time"], "incidence": data},
                                       n_particles=1, n_groups=3)
    names = ["N", "I0", "beta", "gamma"]
    ll = gpu_lib.dust_likelihood_run(obj, [dict(zip(names, p)) for p in pars],
                                     save_trajectories=True)
    h = gpu_lib.dust_likelihood_last_trajectories(obj)
    assert h.shape == (365, 3, 1, 5)
    ref = np.array(g["state"])[:3]                     # (3, len(state_index), 5)
    np.testing.assert_allclose(h[g["state_index"], :, 0, :].transpose(1, 0, 2), ref,
                               rtol=1e-10, atol=1e-9)
    ll2 = gpu_lib.dust_likelihood_run(obj, None, save_trajectories=True, adjoint=True)
    np.testing.assert_array_equal(ll, ll2)
    np.testing.assert_array_equal(gpu_lib.dust_likelihood_last_trajectories(obj), h)
