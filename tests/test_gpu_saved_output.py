"""save_snapshots and index_state as arguments of the run entry points
(dust2_filter_run<T>, dust2_unfilter_run<T>: dust2/r/filter.hpp:21-52,
dust2/r/unfilter.hpp:21-62) and the snapshot exports of every filter and
unfilter (dust2/r/filter.hpp:97-139, dust2/r/unfilter.hpp:97-128;
registered per model by inst/template/compare.cpp:22).  Trajectories keep
the states of index_state at every data time, snapshots the whole state at
their times (dust2/trajectories.hpp:31-57).

Bars: the discrete filter's trajectories and snapshots bit-exact with the
oracle's lineage over the oracle's own saved states and ancestry; the
discrete unfilter's within 1e-12 relative of the oracle's deterministic run
(the GPU fuses multiply-adds: measured differences ~1e-15); the ODE
unfilter's within 1e-10 relative (1e-9 absolute, the bar of its final-state
test) of the oracle solver stopped at each data time; the general-engine filter's last snapshot within 1e-9 of the golden
final state of the reference-built filter, and its unfilter's rows equal to
its own trajectories.  Every call goes through the C ABI."""
import numpy as np
import pytest

import oracle_lib as ol

pytestmark = pytest.mark.gpu

NAMES = ["N", "I0", "beta", "gamma", "exp_noise"]


def _args(snaps, states):
    from dust2_amd import _lib
    s = None if snaps is None else np.ascontiguousarray(snaps, dtype=np.float64)
    st = None if states is None else np.ascontiguousarray(states, dtype=np.uintp)
    keep = (s, st)
    return keep, (_lib.dptr(s), 0 if s is None else s.size, _lib.szptr(st),
                  0 if st is None else st.size)


def _idx(groups):
    return np.ascontiguousarray(groups, dtype=np.uintp)


def test_filter_run_takes_snapshots_and_index_state(gpu_lib):
    from dust2_amd._lib import call, dptr, szptr
    G, P, T = 3, 96, 12
    pars = np.array([[1e4, 10, 0.3 + 0.05 * g, 0.1, 1e6] for g in range(G)])
    time = np.arange(1.0, T + 1)
    data = np.random.default_rng(5).poisson(20, (T, G)).astype(float)
    data[4] = np.nan                     # a time without resampling
    obj = gpu_lib.dust_filter_create(gpu_lib.sir(), 0.0, {"time": time, "incidence": data}, P,
                                     n_groups=G, dt=0.25, seed=9, precision="fp64")
    plist = [dict(zip(NAMES, r)) for r in pars]
    gpu_lib.dust_likelihood_run(obj, plist)
    rng0 = gpu_lib.dust_likelihood_rng_state(obj).reshape(-1, 4)
    snaps = np.array([2.0, 5.0, 12.0])
    states = [4, 1]
    groups = _idx([2, 0])
    keep, save = _args(snaps, states)
    ll = np.zeros(2)
    call("dust2_filter_run", obj.ptr, None, 1, *save, szptr(groups), 2, dptr(ll))
    ref = ol.sir_filter(pars, P, 0.0, time, 0.25, data, rng=rng0, index_group=[2, 0],
                        ancestry=True, traj=True)
    np.testing.assert_allclose(ll, ref["ll"][[2, 0]], rtol=1e-12)
    lineage = ol.trajectory_lineage(ref["traj"], ref["ancestry"])
    traj = np.zeros((T, 2, P, 2))
    call("dust2_filter_last_trajectories", obj.ptr, 0, dptr(traj))
    assert np.array_equal(traj, lineage[:, [2, 0]][..., states])
    snap = np.zeros((3, 2, P, 5))
    call("dust2_filter_last_snapshots", obj.ptr, 0, dptr(snap))
    assert np.array_equal(snap, lineage[[1, 4, 11]][:, [2, 0]])
    # the chosen particle (floor(u P), u from filter stream 2g + 1 after the
    # run) exports the same states along its lineage
    one = np.zeros((T, 2, 1, 2))
    call("dust2_filter_last_trajectories", obj.ptr, 1, dptr(one))
    for j, g in enumerate([2, 0]):
        u, _ = ol.random_real(ref["rng"][g * (P + 2) + 1], 1)
        k = int(np.floor(u[0] * P))
        assert np.array_equal(one[:, j, 0], lineage[:, g, k][..., states])
    # the Python mirror passes the same arguments through
    gpu_lib.dust_likelihood_set_rng_state(obj, rng0)
    ll2 = gpu_lib.dust_likelihood_run(obj, None, save_trajectories=True, save_snapshots=snaps,
                                      index_state=states, index_group=[2, 0])
    np.testing.assert_array_equal(ll2, ll)
    assert np.array_equal(gpu_lib.dust_likelihood_last_trajectories(obj), traj)
    assert np.array_equal(gpu_lib.dust_likelihood_last_snapshots(obj), snap)


def test_filter_run_argument_errors(gpu_lib):
    from dust2_amd._lib import call, dptr
    from dust2_amd._lib import DustError
    time = np.arange(1.0, 6.0)
    obj = gpu_lib.dust_filter_create(gpu_lib.sir(), 0.0, {"time": time, "incidence": np.ones(5)},
                                     16, dt=1.0, seed=1)
    gpu_lib.dust_likelihood_run(obj, dict(zip(NAMES, [1e4, 10, 0.3, 0.1, 1e6])),
                                save_trajectories=True)
    ll = np.zeros(1)
    cases = [([2.0, 2.0], None, "must be increasing"),
             ([2.5], None, "must be found in 'time'"),
             (None, [5], r"'index_state': values must lie in \[0, 4\]"),
             (None, [1, 1], "'index_state' must be distinct")]
    for snaps, states, msg in cases:
        keep, save = _args(snaps, states)
        with pytest.raises(DustError, match=msg):
            call("dust2_filter_run", obj.ptr, None, 1, *save, None, 0, dptr(ll))
        # a refused run leaves the last run's outputs as they were
        assert gpu_lib.dust_likelihood_last_trajectories(obj).shape == (5, 1, 16, 5)
    # a non-null index_state of length zero is refused, as check_index does
    empty = np.zeros(0, dtype=np.uintp)
    from dust2_amd._lib import szptr
    with pytest.raises(DustError, match="nonzero length"):
        call("dust2_filter_run", obj.ptr, None, 1, None, 0, szptr(empty), 0, None, 0, dptr(ll))


def test_discrete_unfilter_snapshots_and_index_state(gpu_lib):
    from dust2_amd._lib import call, dptr, szptr
    G, T = 3, 15
    pars = np.array([[1000, 10, 0.2 + 0.05 * g, 0.1, 1e6] for g in range(G)])
    time = np.arange(4.0, 4.0 * (T + 1), 4.0)
    data = np.random.default_rng(2).poisson(15, (T, G)).astype(float)
    data[3, 1] = np.nan
    obj = gpu_lib.dust_unfilter_create(gpu_lib.sir(), 0.0, {"time": time, "incidence": data},
                                       n_groups=G, dt=0.25)
    gpu_lib.dust_likelihood_run(obj, [dict(zip(NAMES, r)) for r in pars])
    ref = ol.discrete_unfilter(pars, 1, 0.0, time, 0.25, data, trajectories=True)
    snaps = time[[0, 6, 14]]
    keep, save = _args(snaps, [3, 0, 2])
    groups = _idx([1, 2])
    ll = np.zeros((2, 1))
    # arguments in the reference's order: initial, save_trajectories,
    # save_snapshots, adjoint, index_state, index_group
    s, n_s, st, n_st = save
    call("dust2_unfilter_sir_run", obj.ptr, None, 1, s, n_s, 1, st, n_st, szptr(groups), 2,
         dptr(ll))
    np.testing.assert_allclose(ll[:, 0], ref["ll"][[1, 2]], rtol=1e-12)
    traj = np.zeros((T, 2, 1, 3))
    call("dust2_unfilter_sir_last_trajectories", obj.ptr, dptr(traj))
    np.testing.assert_allclose(traj, ref["traj"][:, [1, 2]][..., [3, 0, 2]], rtol=1e-12)
    snap = np.zeros((3, 2, 1, 5))
    call("dust2_unfilter_sir_last_snapshots", obj.ptr, dptr(snap))
    np.testing.assert_allclose(snap, ref["traj"][[0, 6, 14]][:, [1, 2]], rtol=1e-12)
    # both read the same saved states: the kept columns agree exactly
    np.testing.assert_array_equal(snap[..., [3, 0, 2]], traj[[0, 6, 14]])
    # the adjoint of the same run: the reference's form, NaN where data are
    # missing (group 1)
    grad = np.zeros((2, 3))
    call("dust2_unfilter_sir_last_gradient", obj.ptr, dptr(grad))
    _, ref_gr = ol.sir_unfilter(pars, 0.0, time, 0.25, data)
    np.testing.assert_allclose(grad, ref_gr[[1, 2]], rtol=1e-10)
    # snapshots without trajectories, then a run without either
    gpu_lib.dust_likelihood_run(obj, None, save_snapshots=[time[2]])
    with pytest.raises(gpu_lib.DustError, match="Trajectories are not current"):
        gpu_lib.dust_likelihood_last_trajectories(obj)
    np.testing.assert_allclose(gpu_lib.dust_likelihood_last_snapshots(obj), ref["traj"][[2]],
                               rtol=1e-12)
    gpu_lib.dust_likelihood_run(obj, None)
    with pytest.raises(gpu_lib.DustError, match="Snapshots are not current"):
        gpu_lib.dust_likelihood_last_snapshots(obj)


@pytest.mark.parametrize("adjoint", [False, True])
def test_ode_unfilter_snapshots_and_index_state(gpu_lib, adjoint):
    from dust2_amd._lib import call, dptr, szptr
    G, P = 2, 3
    pars = np.array([[1000, 10, 0.25, 0.1], [2000, 5, 0.4, 0.15]])
    time = np.arange(5.0, 65.0, 5.0)
    T = time.size
    data = np.random.default_rng(4).poisson(30, (T, G)).astype(float)
    obj = gpu_lib.dust_unfilter_create(gpu_lib.sirode(), 0.0, {"time": time, "incidence": data},
                                       n_particles=P, n_groups=G)
    gpu_lib.dust_likelihood_run(obj, [dict(zip(["N", "I0", "beta", "gamma"], r))
                                      for r in pars])
    snaps = time[[1, 7, 11]]
    keep, save = _args(snaps, [2, 4])
    ll = np.zeros((G, P))
    name = "dust2_unfilter_sirode_run_gradient" if adjoint else "dust2_unfilter_sirode_run"
    call(name, obj.ptr, 1, *save, None, 0, dptr(ll))
    ref = ol.sirode_unfilter(pars, P, 0.0, time, data)
    np.testing.assert_allclose(ll.reshape(-1), ref["ll"], rtol=1e-10)
    traj = np.zeros((T, G, P, 2))
    call("dust2_unfilter_sirode_last_trajectories", obj.ptr, dptr(traj))
    snap = np.zeros((3, G, P, 5))
    call("dust2_unfilter_sirode_last_snapshots", obj.ptr, dptr(snap))
    # the oracle solver stopped at each data time gives that row
    for row, t in enumerate(time):
        at = ol.sirode_unfilter(pars, P, 0.0, time[:row + 1], data[:row + 1])["state"]
        np.testing.assert_allclose(traj[row], at[..., [2, 4]], rtol=1e-10, atol=1e-9)
        if t in snaps:
            np.testing.assert_allclose(snap[list(snaps).index(t)], at, rtol=1e-10, atol=1e-9)
    # one group: the rows of that group only
    one = _idx([1])
    ll1 = np.zeros((1, P))
    keep, save = _args(snaps[:1], None)
    call("dust2_unfilter_sirode_run", obj.ptr, 0, *save, szptr(one), 1, dptr(ll1))
    s1 = np.zeros((1, 1, P, 5))
    call("dust2_unfilter_sirode_last_snapshots", obj.ptr, dptr(s1))
    np.testing.assert_array_equal(s1[0], snap[0][[1]])
    with pytest.raises(gpu_lib.DustError, match="Trajectories are not current"):
        call("dust2_unfilter_sirode_last_trajectories", obj.ptr, dptr(np.zeros(T * P * 5)))


def test_general_engine_filter_and_unfilter_snapshots(gpu_lib):
    import json
    import os
    from dust2_amd._lib import call, dptr
    path = os.path.join(os.path.dirname(__file__), "golden", "malaria_filter.json")
    with open(path) as f:
        case = json.load(f)["cases"][0]
    P = case["P"]
    d = np.array([[[np.nan if v is None else v for v in f] for f in g] for g in case["data"]],
                 dtype=float)
    times = np.array(case["times"])
    data = {"time": times, "tested": d[:, :, 0], "positive": d[:, :, 1]}
    obj = gpu_lib.dust_filter_create(gpu_lib.malaria(), 0.0, data, P, dt=case["dt"] or None,
                                     seed=np.array(case["streams"], dtype=np.uint64))
    gpu_lib.dust_likelihood_ensure_initialised(obj, case["pars"][0])
    T = times.size
    keep, save = _args(times[[0, T - 1]], [1, 0])
    ll = np.zeros(1)
    call("dust2_filter_ode_run", obj.ptr, None, 1, *save, None, 0, dptr(ll))
    np.testing.assert_allclose(ll, case["runs"][0]["ll"], rtol=1e-10)
    traj = np.zeros((T, 1, P, 2))
    call("dust2_filter_ode_last_trajectories", obj.ptr, 0, dptr(traj))
    snap = np.zeros((2, 1, P, 20))
    call("dust2_filter_ode_last_snapshots", obj.ptr, 0, dptr(snap))
    final = np.array(case["runs"][0]["state"])
    np.testing.assert_allclose(snap[1], final, rtol=1e-9, atol=1e-12)
    np.testing.assert_array_equal(traj[-1, ..., 0], snap[1, ..., 1])
    np.testing.assert_array_equal(snap[0][..., [1, 0]], traj[0])

    un = gpu_lib.dust_unfilter_create(gpu_lib.malaria(), 0.0, data, n_particles=2)
    gpu_lib.dust_likelihood_run(un, {"a": 0.4}, save_trajectories=True)
    full = gpu_lib.dust_likelihood_last_trajectories(un)
    keep, save = _args(times[[1, 2]], [5, 1])
    ll2 = np.zeros((1, 2))
    call("dust2_unfilter_ode_run", un.ptr, None, 1, *save, None, 0, dptr(ll2))
    t2 = np.zeros((T, 1, 2, 2))
    call("dust2_unfilter_ode_last_trajectories", un.ptr, dptr(t2))
    np.testing.assert_array_equal(t2, full[..., [5, 1]])
    s2 = np.zeros((2, 1, 2, 20))
    call("dust2_unfilter_ode_last_snapshots", un.ptr, dptr(s2))
    np.testing.assert_array_equal(s2, full[[1, 2]])
