"""The SIR with beta interpolated in time on the GPU (pipeline engine)
against the oracle's sir_tv model (pinned in test_oracle_interpolate.py):
FP64 draw for draw for every interpolation kind, a flat schedule equal to
the plain SIR, parameter updates of the schedule, and the reference's
error for a step time outside the knots."""
import numpy as np
import pytest

import oracle_lib as ol

pytestmark = pytest.mark.gpu

NAMES = ["N", "I0", "gamma", "exp_noise"]


def sir_data(n_time, seed):
    rng = ol.prng(1, seed=seed)
    out, _ = ol.sir_simulate(np.array([[1000.0, 10.0, 0.25, 0.1, 1e6]]), 1, 0.0, 0.25,
                             np.arange(1.0, n_time + 1), rng)
    return out[:, 0, 0, 4].copy()


def run(lib, kind, pars, beta_t, beta_y, P, time, data, seed, precision="fp64"):
    G = len(pars)
    obj = lib.dust_filter_create(lib.sir_time_varying(kind), 0.0,
                                 {"time": time, "incidence": data}, P, n_groups=G, dt=0.25,
                                 seed=seed, precision=precision)
    plist = [dict(zip(NAMES, p[[0, 1, 3, 4]]), beta_t=beta_t, beta_y=beta_y[g])
             for g, p in enumerate(pars)]
    ll = lib.dust_likelihood_run(obj, plist, save_trajectories=True)
    return obj, np.atleast_1d(ll)


@pytest.mark.parametrize("kind", ["constant", "linear", "spline"])
def test_fp64_matches_oracle(gpu_lib, kind):
    G, P, n_time = 3, 200, 40
    pars = np.array([[1000.0, 10.0, 0.0, 0.1, 1e6], [2000.0, 5.0, 0.0, 0.15, 1e6],
                     [1500.0, 20.0, 0.0, 0.08, 1e6]])
    beta_t = np.array([0.0, 5.0, 12.0, 20.0, 33.0, 45.0])
    beta_y = np.array([[0.3, 0.5, 0.2, 0.25, 0.1, 0.3],
                       [0.2, 0.2, 0.4, 0.3, 0.2, 0.1],
                       [0.6, 0.3, 0.3, 0.1, 0.2, 0.2]])
    time = np.arange(1.0, n_time + 1)
    data = np.stack([sir_data(n_time, 40 + g) for g in range(G)], axis=1)
    obj, ll = run(gpu_lib, kind, pars, beta_t, beta_y, P, time, data, 7)
    ref = ol.sir_filter(ol.sir_tv_rows(pars, kind, beta_t, beta_y), P, 0.0, time, 0.25, data,
                        seed=7, model="sir_tv", ancestry=True, n_threads=4)
    np.testing.assert_allclose(ll, ref["ll"], rtol=1e-12)
    assert np.array_equal(gpu_lib.dust_likelihood_last_state(obj), ref["state"])
    assert np.array_equal(gpu_lib.dust_likelihood_last_ancestry(obj), ref["ancestry"])
    assert gpu_lib.dust_likelihood_last_stats(obj)["engine"] == "pipeline"
    # FP32 runs the same schedule
    _, ll32 = run(gpu_lib, kind, pars, beta_t, beta_y, P, time, data, 7, precision="fp32")
    assert np.all(np.isfinite(ll32))


def test_flat_schedule_equals_plain_sir(gpu_lib):
    P, time = 300, np.arange(1.0, 31.0)
    data = sir_data(30, 3)
    pars = np.array([[1000.0, 10.0, 0.0, 0.1, 1e6]])
    obj, ll = run(gpu_lib, "linear", pars, np.array([0.0, 10.0, 31.0]),
                  np.full((1, 3), 0.25), P, time, data, 11)
    plain = gpu_lib.dust_filter_create(gpu_lib.sir(), 0.0, {"time": time, "incidence": data}, P,
                                       dt=0.25, seed=11)
    plain.set_engine("pipeline")
    llp = gpu_lib.dust_likelihood_run(plain, {"N": 1000.0, "I0": 10.0, "beta": 0.25,
                                              "gamma": 0.1, "exp_noise": 1e6})
    np.testing.assert_array_equal(ll, [llp])
    np.testing.assert_array_equal(gpu_lib.dust_likelihood_last_state(obj),
                                  gpu_lib.dust_likelihood_last_state(plain))


def test_schedule_updates_and_errors(gpu_lib):
    P, time = 100, np.arange(1.0, 21.0)
    data = sir_data(20, 5)
    pars = np.array([[1000.0, 10.0, 0.0, 0.1, 1e6]] * 2)
    beta_t = np.array([0.0, 10.0, 25.0])
    beta_y = np.array([[0.3, 0.1, 0.2], [0.2, 0.4, 0.1]])
    obj, ll = run(gpu_lib, "linear", pars, beta_t, beta_y, P, time, data, 13)
    # a new schedule for group 1 only, through index_group
    new_y = np.array([0.5, 0.5, 0.5])
    rng = gpu_lib.dust_likelihood_rng_state(obj)
    gpu_lib.dust_likelihood_run(obj, [{"beta_t": beta_t, "beta_y": new_y}], index_group=[1])
    gpu_lib.dust_likelihood_set_rng_state(obj, rng)
    ll2 = gpu_lib.dust_likelihood_run(obj, None)
    ref = ol.sir_filter(ol.sir_tv_rows(pars, "linear", beta_t, np.stack([beta_y[0], new_y])),
                        P, 0.0, time, 0.25, data[:, None], rng=rng.reshape(-1, 4), n_threads=4,
                        model="sir_tv")
    np.testing.assert_allclose(ll2, ref["ll"], rtol=1e-12)
    # the persistent engine does not take an interpolated beta
    obj.set_engine("persistent")
    with pytest.raises(gpu_lib.DustError, match="pipeline engine only"):
        gpu_lib.dust_likelihood_run(obj, None)
    # steps beyond the last knot: every particle reports the reference's error
    short = np.array([0.0, 9.0])
    with pytest.raises(gpu_lib.DustError) as e:
        run(gpu_lib, "linear", pars[:1], short, np.array([[0.2, 0.3]]), P, time, data, 1)
    msg = str(e.value)
    assert msg.startswith(f"{P} particles reported errors.")
    assert "  - 1: Tried to interpolate at time = 9.25, which is 0.25 after the last time (9)" \
        in msg
    with pytest.raises(ValueError, match="'beta_t' and 'beta_y' are required"):
        run(gpu_lib, "linear", pars[:1], None, [None], P, time, data, 1)
