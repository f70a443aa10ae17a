"""Groups larger than one workgroup's LDS (above 16,384 particles) resample
across workgroups (resample_kernels.hip).  Bars: ancestors bit-exact with
the oracle's restatement of scale_log_weights / resample_weight
(dust2/filter_details.hpp:11-74) for identical weights, including the
sequential fallback forced by ties, NaN and -Inf weights; the
log-likelihood increment within 1e-10 absolute (a sum of up to 10^6 terms
in another order); the FP64 filter at 20,000 particles per group identical
to the oracle draw for draw."""
import numpy as np
import pytest

import oracle_lib as ol

pytestmark = pytest.mark.gpu


def gpu_resample(lib, w, u):
    from dust2_amd import _lib
    w = np.ascontiguousarray(w, dtype=np.float64)
    idx = np.zeros(w.size, dtype=np.uint32)
    out3 = np.zeros(3)
    _lib.call("dust2_amd_resample", _lib.dptr(w), w.size, float(u), _lib.u32ptr(idx),
              _lib.dptr(out3))
    return idx, out3


@pytest.mark.parametrize("n,u,kind", [(100000, 0.3, "normal"), (16385, 0.999, "normal"),
                                      (250000, 0.0, "spread"), (50000, 0.0, "ties"),
                                      (40000, 0.7, "nan_inf"), (1 << 20, 0.123, "normal")])
def test_multi_workgroup_resample_bit_exact(gpu_lib, n, u, kind):
    rng = np.random.default_rng(n)
    if kind == "normal":
        w = rng.normal(-40, 4, n)
    elif kind == "spread":
        w = rng.uniform(-700, 0, n)
    elif kind == "ties":
        w = np.zeros(n) - 1.0          # systematic points on the partial sums
    else:
        w = rng.normal(-3, 1, n)
        w[::7] = -np.inf
        w[3::101] = np.nan
    idx, out3 = gpu_resample(gpu_lib, w, u)
    inc, we = ol.scale_log_weights(w)
    expect = ol.resample_weight(we, u)
    assert np.array_equal(idx, expect)
    assert out3[1] == 1
    assert abs(out3[0] - inc) <= 1e-10 * max(1.0, abs(inc))
    if kind == "ties":
        assert out3[2] == 1            # decided by the sequential fallback


def test_multi_workgroup_special_weights(gpu_lib):
    # all zero: identity, nothing drawn; all impossible: -Inf and ancestor 0
    idx, out3 = gpu_resample(gpu_lib, np.zeros(30000), 0.4)
    assert np.array_equal(idx, np.arange(30000)) and out3[1] == 0 and out3[0] == 0
    idx, out3 = gpu_resample(gpu_lib, np.full(30000, -np.inf), 0.4)
    assert np.all(idx == 0) and out3[0] == -np.inf


@pytest.mark.parametrize("precision", ["fp64"])
def test_filter_beyond_one_workgroup_matches_oracle(gpu_lib, precision):
    G, P = 2, 20000
    pars = np.array([[5e5, 10, 0.45, 0.1, 1e6], [2e5, 20, 0.35, 0.12, 1e6]])
    time = np.arange(1.0, 13.0)
    sim, _ = ol.sir_simulate(pars, 1, 0.0, 0.25, time, ol.prng(G, seed=3))
    data = sim[:, :, 0, 4].copy()
    data[4, 1] = np.nan
    names = ["N", "I0", "beta", "gamma", "exp_noise"]
    obj = gpu_lib.dust_filter_create(gpu_lib.sir(), 0.0, {"time": time, "incidence": data}, P,
                                     n_groups=G, dt=0.25, seed=11, precision=precision)
    ll = gpu_lib.dust_likelihood_run(obj, [dict(zip(names, r)) for r in pars],
                                     save_trajectories=True)
    ref = ol.sir_filter(pars, P, 0.0, time, 0.25, data, seed=11, ancestry=True, n_threads=8)
    np.testing.assert_allclose(ll, ref["ll"], rtol=1e-12)
    assert np.array_equal(gpu_lib.dust_likelihood_last_state(obj), ref["state"])
    assert np.array_equal(gpu_lib.dust_likelihood_last_ancestry(obj), ref["ancestry"])
    assert np.array_equal(gpu_lib.dust_likelihood_rng_state(obj).reshape(-1, 4), ref["rng"])
