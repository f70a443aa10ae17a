"""Boundary behaviours of the filter and server entry points: groups left
out of index_group keep the caller's initial state, repeated groups are
refused at the C ABI, trajectories of a failed run are not current, and
served ODE requests with different start times each match the plain
unfilter."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SIR = {"N": 1000.0, "I0": 10.0, "beta": 0.2, "gamma": 0.1, "exp_noise": 1e6}


@pytest.mark.parametrize("engine", ["persistent", "pipeline"])
@pytest.mark.parametrize("n_time", [4, 5])
def test_unselected_groups_keep_initial_state(gpu_lib, engine, n_time):
    # dust2/r/filter.hpp:21-52 sets the state of every group and runs only
    # index_group; the others keep what the caller gave
    time = np.arange(1.0, n_time + 1)
    data = np.random.default_rng(1).poisson(5, (n_time, 2)).astype(float)
    obj = gpu_lib.dust_filter_create(gpu_lib.sir(), 0.0, {"time": time, "incidence": data},
                                     64, n_groups=2, dt=0.25, seed=3)
    obj.set_engine(engine)
    init = np.zeros((2, 64, 5))
    init[:, :, 0] = 990.0
    init[:, :, 1] = 10.0
    init[1, :, 2] = np.arange(64)        # marks group 1
    gpu_lib.dust_likelihood_run(obj, [SIR, SIR])
    gpu_lib.dust_likelihood_run(obj, None, initial=init, index_group=[0])
    st = gpu_lib.dust_likelihood_last_state(obj)
    np.testing.assert_array_equal(st[1], init[1])
    assert not np.array_equal(st[0], init[0])


def test_repeated_index_group_refused_at_the_c_abi(gpu_lib):
    from dust2_amd import _lib
    time = np.arange(1.0, 4.0)
    data = np.ones((3, 2))
    obj = gpu_lib.dust_filter_create(gpu_lib.sir(), 0.0, {"time": time, "incidence": data},
                                     32, n_groups=2, dt=1.0, seed=1)
    gpu_lib.dust_likelihood_run(obj, [SIR, SIR])
    idx = np.array([1, 1], dtype=np.uintp)
    ll = np.zeros(2)
    buf = C.create_string_buffer(1024)
    rc = _lib.lib().dust2_filter_run(obj.ptr, None, 0, None, 0, None, 0, _lib.szptr(idx), 2,
                                     _lib.dptr(ll), buf, len(buf))
    assert rc != 0
    assert "must be distinct" in buf.value.decode()


def test_trajectories_of_failed_run_are_not_current(gpu_lib):
    time = np.arange(1.0, 6.0)
    data = np.ones(5)
    obj = gpu_lib.dust_filter_create(gpu_lib.sir(), 0.0, {"time": time, "incidence": data},
                                     32, dt=1.0, seed=1)
    gpu_lib.dust_likelihood_run(obj, SIR, save_trajectories=True)
    assert gpu_lib.dust_likelihood_last_trajectories(obj).shape == (5, 1, 32, 5)
    with pytest.raises(gpu_lib.DustError, match="particles reported errors"):
        gpu_lib.dust_likelihood_run(obj, dict(SIR, gamma=-1.0), save_trajectories=True)
    with pytest.raises(gpu_lib.DustError, match="not current"):
        gpu_lib.dust_likelihood_last_trajectories(obj)


def test_served_ode_requests_keep_their_own_start_time(gpu_lib):
    from dust2_amd.serving import Server
    rng = np.random.default_rng(8)
    reqs = []
    for i, t0 in enumerate([0.0, 3.0, 10.0, 1.0]):
        pars = {"N": 1000.0, "I0": 10.0, "beta": rng.uniform(0.2, 0.5),
                "gamma": rng.uniform(0.05, 0.2)}
        time = t0 + np.arange(1.0, 41.0)
        data = {"time": time, "incidence": rng.poisson(6, time.size).astype(float)}
        reqs.append((t0, pars, data))
    sv = Server(hbm_bytes=1 << 26, chunk_intervals=8)
    ids = [sv.submit(gpu_lib.sirode(), p, 1, t0, d, dt=1.0) for (t0, p, d) in reqs]
    res = {r[0]: r for r in sv.drain()}
    for rid, (t0, p, d) in zip(ids, reqs):
        obj = gpu_lib.dust_unfilter_create(gpu_lib.sirode(), t0, d, n_particles=1)
        ll = obj.run(p)
        assert res[rid][2] == 0
        assert res[rid][1] == ll, (t0, res[rid][1], ll)
