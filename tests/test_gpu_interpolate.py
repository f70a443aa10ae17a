"""Device interpolation of time-varying inputs against the oracle (itself
pinned by the reference's interpolators, closed forms and scipy's natural
cubic spline in test_oracle_interpolate.py): identical values, and the
reference's error texts (tests/testthat/test-interpolate.R)."""
import numpy as np
import pytest

import oracle_lib as ol

pytestmark = pytest.mark.gpu

KINDS = ["constant", "linear", "spline"]


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("n_series", [1, 4])
def test_device_equals_oracle(gpu_lib, kind, n_series):
    rng = np.random.default_rng(10 + n_series)
    t = np.cumsum(rng.uniform(0.05, 2.0, 200))
    y = rng.normal(size=(200, n_series)) if n_series > 1 else rng.normal(size=200)
    z = np.concatenate([t, rng.uniform(t[0], t[-1], 20000)])
    np.testing.assert_array_equal(gpu_lib.gpu_interpolate(kind, t, y, z),
                                  ol.interpolate(kind, t, y, z))


def test_search_positions(gpu_lib):
    # test-interpolate.R:1-29
    t = np.arange(10) + 0.5
    eps = 1e-7
    z = [0.5, 0.5 + eps, 1.5 - eps, 1.5, 1.5 + eps, 2.5 - eps, 5.5, 5.5 + eps, 6.5 - eps,
         9.5 - eps, 9.5, 100.0]
    got = gpu_lib.gpu_interpolate("constant", t, np.arange(10.0), z)
    np.testing.assert_array_equal(got, [0, 0, 0, 1, 1, 1, 5, 5, 5, 8, 9, 9])


def test_errors(gpu_lib):
    t = np.arange(10) + 0.5
    y = np.arange(10.0)
    with pytest.raises(gpu_lib.DustError) as e:
        gpu_lib.gpu_interpolate("constant", t, y, [0.0])
    assert str(e.value) == ("Tried to interpolate at time = 0, which is 0.5 before the "
                            "first time (0.5)")
    for kind in ["linear", "spline"]:
        with pytest.raises(gpu_lib.DustError,
                           match=r"Tried to interpolate at time = .+, which is .+ after the "
                                 r"last time \(9.5\)"):
            gpu_lib.gpu_interpolate(kind, t, y, [9.5 + 1e-7])
    with pytest.raises(gpu_lib.DustError, match="before the first time"):
        gpu_lib.gpu_interpolate("linear", t, y, [np.nan])
    # dust2/interpolate/interpolate.hpp:48-66
    with pytest.raises(gpu_lib.DustError,
                       match="Time variable 't' and interpolation target 'y' must have the same "
                             r"length, but do not \(10 vs 9\)"):
        gpu_lib.gpu_interpolate("linear", t, y[:9], [1.0])
    bad = t.copy()
    bad[4] = bad[3]
    with pytest.raises(gpu_lib.DustError,
                       match="Time variable 't' must be strictly increasing but was not at "
                             "index 4"):
        gpu_lib.gpu_interpolate("spline", bad, y, [1.0])
