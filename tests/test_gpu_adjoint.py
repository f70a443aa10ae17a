"""Deterministic discrete-time SIR unfilter and its adjoint gradient on the
GPU against the FP64 oracle, in both forms: the reference's hand-written
adjoint (the default; pinned in test_oracle_ode.py by finite differences of
the quantities it actually differentiates) and the exact derivative (pinned
by central finite differences of the likelihood); and the reference's own
adjoint tests (tests/testthat/test-unfilter-adjoint.R).  Bar: 1e-10
relative."""
import numpy as np
import pytest

import oracle_lib as ol

pytestmark = pytest.mark.gpu

NAMES = ["N", "I0", "beta", "gamma", "exp_noise"]


def plist(pars):
    return [dict(zip(NAMES, p)) for p in pars]


def make(lib, pars, time, dt, data, P=1, form="reference"):
    obj = lib.dust_unfilter_create(lib.sir(), 0.0, {"time": time, "incidence": data},
                                   n_particles=P, n_groups=len(pars), dt=dt, adjoint_form=form)
    return obj


def random_pars(n, seed):
    rng = np.random.default_rng(seed)
    return np.column_stack([np.round(rng.uniform(1e3, 1e5, n)), rng.uniform(5, 50, n),
                            rng.uniform(0.1, 0.5, n), rng.uniform(0.05, 0.2, n),
                            np.full(n, 1e6)])


@pytest.mark.parametrize("form", ["reference", "exact"])
@pytest.mark.parametrize("times,dt", [([4.0, 8.0, 12.0, 16.0], 1.0),
                                      ([3.0, 6.0, 9.0, 12.0], 1.0),
                                      ([4.0, 8.0, 12.0, 16.0], 0.25)])
def test_reference_adjoint_cases(gpu_lib, times, dt, form):
    pars = np.array([[1000, 10, 0.1, 0.2, 1e6]], dtype=float)
    time = np.array(times)
    data = np.arange(1.0, 5.0)
    obj = make(gpu_lib, pars, time, dt, data, form=form)
    ll1 = gpu_lib.dust_likelihood_run(obj, plist(pars)[0], adjoint=False)
    ll2 = gpu_lib.dust_likelihood_run(obj, plist(pars)[0], adjoint=True)
    assert ll1 == ll2                      # test-unfilter-adjoint.R:1-11
    gr = gpu_lib.dust_likelihood_last_gradient(obj)
    ref_ll, ref_gr = ol.sir_unfilter(pars, 0.0, time, dt, data[:, None], exact=form == "exact")
    assert ll2 == pytest.approx(ref_ll[0], rel=1e-12)
    np.testing.assert_allclose(gr, ref_gr[0], rtol=1e-10)


def test_reference_form_passes_the_reference_test(gpu_lib):
    # test-unfilter-adjoint.R:14-27 compares with numDeriv through
    # expect_equal(tolerance = 1e-4), i.e. the mean relative difference
    # sum|a - b| / sum|a| of the whole vector; the reference's gradient
    # passes it although its I0 component is not the derivative
    pars = np.array([[1000, 10, 0.1, 0.2, 1e6]], dtype=float)
    time = np.array([4.0, 8.0, 12.0, 16.0])
    data = np.arange(1.0, 5.0)
    obj = make(gpu_lib, pars, time, 1.0, data)
    gpu_lib.dust_likelihood_run(obj, plist(pars)[0], adjoint=True)
    gr = gpu_lib.dust_likelihood_last_gradient(obj)
    _, exact = ol.sir_unfilter(pars, 0.0, time, 1.0, data[:, None], exact=True)
    assert np.sum(np.abs(exact[0] - gr)) / np.sum(np.abs(exact[0])) < 1e-4
    assert abs(gr[2] - exact[0, 2]) > 1e-3 * abs(exact[0, 2])


@pytest.mark.parametrize("form", ["reference", "exact"])
def test_many_groups_match_oracle(gpu_lib, form):
    G = 2048
    pars = random_pars(G, 1)
    time = np.arange(1.0, 101.0)
    rng = np.random.default_rng(2)
    data = rng.poisson(30, (100, G)).astype(float)
    data[rng.random((100, G)) < 0.05] = np.nan
    data[:, :G // 2] = np.nan_to_num(data[:, :G // 2], nan=7.0)   # half without gaps
    obj = make(gpu_lib, pars, time, 0.25, data, form=form)
    ll = gpu_lib.dust_likelihood_run(obj, plist(pars), adjoint=True)
    gr = gpu_lib.dust_likelihood_last_gradient(obj)
    ref_ll, ref_gr = ol.sir_unfilter(pars, 0.0, time, 0.25, data, exact=form == "exact")
    np.testing.assert_allclose(ll, ref_ll, rtol=1e-11)
    # the reference's form has NaN gradients where data are missing
    assert np.array_equal(np.isnan(gr), np.isnan(ref_gr))
    assert np.isnan(gr).any() == (form == "reference")
    ok = ~np.isnan(ref_gr)
    np.testing.assert_allclose(gr[ok], ref_gr[ok], rtol=1e-10,
                               atol=1e-9 * np.abs(ref_gr[ok]).max())
    # the final state is the deterministic trajectory's end
    st = obj.last_state()
    assert st.shape == (G, 1, 5)
    np.testing.assert_allclose(st[:, 0, 0] + st[:, 0, 1] + st[:, 0, 2], pars[:, 0], rtol=1e-12)


def test_gradient_matches_central_differences_on_gpu(gpu_lib):
    pars = random_pars(16, 5)
    time = np.arange(1.0, 31.0)
    data = np.random.default_rng(6).poisson(15, (30, 16)).astype(float)
    obj = make(gpu_lib, pars, time, 0.5, data, form="exact")
    gpu_lib.dust_likelihood_run(obj, plist(pars), adjoint=True)
    gr = gpu_lib.dust_likelihood_last_gradient(obj)
    fd = np.zeros_like(gr)
    for j, col in enumerate([2, 3, 1]):
        h = 1e-6 * pars[:, col]
        up, dn = pars.copy(), pars.copy()
        up[:, col] += h
        dn[:, col] -= h
        lu = gpu_lib.dust_likelihood_run(obj, plist(up))
        ld = gpu_lib.dust_likelihood_run(obj, plist(dn))
        fd[:, j] = (lu - ld) / (2 * h)
    np.testing.assert_allclose(gr, fd, rtol=1e-5)


def test_index_group_and_particles(gpu_lib):
    pars = random_pars(6, 7)
    time = np.arange(1.0, 21.0)
    data = np.random.default_rng(8).poisson(10, (20, 6)).astype(float)
    obj = make(gpu_lib, pars, time, 1.0, data, P=3)
    full = gpu_lib.dust_likelihood_run(obj, plist(pars), adjoint=True)
    assert full.shape == (6, 3)
    assert np.all(full == full[:, :1])     # deterministic particles are identical
    with pytest.raises(gpu_lib.DustError, match="n_particles > 1 not yet supported"):
        gpu_lib.dust_likelihood_last_gradient(obj)
    obj1 = make(gpu_lib, pars, time, 1.0, data)
    all_ll = gpu_lib.dust_likelihood_run(obj1, plist(pars), adjoint=True)
    all_gr = gpu_lib.dust_likelihood_last_gradient(obj1)
    sub = gpu_lib.dust_likelihood_run(obj1, None, index_group=[4, 1], adjoint=True)
    np.testing.assert_array_equal(sub, all_ll[[4, 1]])
    np.testing.assert_array_equal(gpu_lib.dust_likelihood_last_gradient(obj1), all_gr[[4, 1]])


def test_gradient_errors_and_initial_state(gpu_lib):
    pars = random_pars(2, 9)
    time = np.arange(1.0, 11.0)
    data = np.full((10, 2), 5.0)
    obj = make(gpu_lib, pars, time, 1.0, data, form="exact")
    gpu_lib.dust_likelihood_run(obj, plist(pars))
    with pytest.raises(gpu_lib.DustError, match="System was not run with 'adjoint = TRUE'"):
        gpu_lib.dust_likelihood_last_gradient(obj)
    # a given initial state does not depend on I0 (exact form); the
    # reference's form applies its adjoint_initial whatever the initial
    # state (dust2/unfilter.hpp:219-222), so its gradient is unchanged
    init = np.zeros((2, 1, 5))
    init[:, 0, 0] = pars[:, 0] - pars[:, 1]
    init[:, 0, 1] = pars[:, 1]
    ll_a = gpu_lib.dust_likelihood_run(obj, None, adjoint=True)
    gr_a = gpu_lib.dust_likelihood_last_gradient(obj)
    ll_b = gpu_lib.dust_likelihood_run(obj, None, initial=init, adjoint=True)
    gr_b = gpu_lib.dust_likelihood_last_gradient(obj)
    np.testing.assert_array_equal(ll_a, ll_b)
    np.testing.assert_array_equal(gr_a[:, :2], gr_b[:, :2])
    assert np.all(gr_b[:, 2] == 0)
    ref = make(gpu_lib, pars, time, 1.0, data)
    gpu_lib.dust_likelihood_run(ref, plist(pars), adjoint=True)
    gr_c = gpu_lib.dust_likelihood_last_gradient(ref)
    gpu_lib.dust_likelihood_run(ref, None, initial=init, adjoint=True)
    np.testing.assert_array_equal(gpu_lib.dust_likelihood_last_gradient(ref), gr_c)
    # invalid parameters report the reference's binomial error
    with pytest.raises(gpu_lib.DustError, match="Invalid call to binomial"):
        gpu_lib.dust_likelihood_run(obj, [{"beta": -0.5}, {"beta": 0.2}])


def test_unfilter_trajectories_with_and_without_adjoint(gpu_lib):
    # test-unfilter-adjoint.R:100-117 and test-unfilter.R:46-75
    pars = random_pars(3, 12)
    time = np.arange(1.0, 31.0)
    data = np.random.default_rng(13).poisson(10, (30, 3)).astype(float)
    obj = make(gpu_lib, pars, time, 0.25, data)
    with pytest.raises(gpu_lib.DustError, match="Trajectories are not current"):
        gpu_lib.dust_likelihood_run(obj, plist(pars))
        gpu_lib.dust_likelihood_last_trajectories(obj)
    ll1 = gpu_lib.dust_likelihood_run(obj, None, save_trajectories=True)
    h1 = gpu_lib.dust_likelihood_last_trajectories(obj)
    ll2 = gpu_lib.dust_likelihood_run(obj, None, save_trajectories=True, adjoint=True)
    h2 = gpu_lib.dust_likelihood_last_trajectories(obj)
    np.testing.assert_array_equal(ll1, ll2)
    np.testing.assert_array_equal(h1, h2)
    assert h1.shape == (30, 3, 1, 5)
    np.testing.assert_array_equal(h1[-1], obj.last_state())
    # the likelihood is the sum of the Poisson terms at the saved states
    from scipy import stats
    expect = stats.poisson.logpmf(data, h1[:, :, 0, 4] + 1e-6).sum(axis=0)
    np.testing.assert_allclose(ll1, expect, rtol=1e-12)
    sub = gpu_lib.dust_likelihood_run(obj, None, save_trajectories=True, index_group=[2])
    np.testing.assert_array_equal(gpu_lib.dust_likelihood_last_trajectories(obj)[:, 0],
                                  h1[:, 2])
    assert sub.shape == (1,)
