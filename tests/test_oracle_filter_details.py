"""Pins the oracle's log-sum-exp and systematic resampling against the
reference's own dust2/filter_details.hpp (golden vectors written by
tests/golden/make_golden.py from oracle/_ref) and against the closed forms of
tests/testthat/test-filter-details.R."""
import json
import os

import numpy as np
import pytest

import oracle_lib as ol

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "filter_details.json")


def load_cases():
    with open(GOLDEN) as f:
        return json.load(f)


def test_golden_vectors_bit_exact():
    for case in load_cases():
        w = np.array(case["w"], dtype=np.float64)
        ll, scaled = ol.scale_log_weights(w)
        assert (ll == case["ll"]) or (np.isnan(ll) and np.isnan(case["ll"]))
        np.testing.assert_array_equal(scaled, np.array(case["scaled"]))
        idx = ol.resample_weight(scaled, case["u"])
        assert idx.tolist() == case["idx"]


@pytest.mark.skipif(not ol.ref_available(), reason="reference build not present")
def test_against_reference_build_random():
    rng = np.random.default_rng(7)
    for n in [2, 5, 31, 500, 2048, 8192]:
        for _ in range(5):
            w = rng.normal(0, rng.choice([0.01, 1, 30]), size=n)
            ll_o, s_o = ol.scale_log_weights(w)
            ll_r, s_r = ol.ref_scale_log_weights(w)
            assert ll_o == ll_r
            np.testing.assert_array_equal(s_o, s_r)
            u = rng.random()
            assert np.array_equal(ol.resample_weight(s_o, u), ol.ref_resample_weight(s_r, u))


def test_resample_closed_form():
    # findInterval(u/n + (0:(n-1))/n, cumsum(w/sum(w))), test-filter-details.R:1-22
    rng = np.random.default_rng(1)
    w = np.abs(rng.standard_cauchy(20))
    for u in [rng.random(), 0.0, 1 - 1e-8]:
        n = w.size
        uu = u / n + np.arange(n) / n
        cw = np.cumsum(w / w.sum())
        expect = np.searchsorted(cw, uu, side="right")
        assert ol.resample_weight(w, u).tolist() == expect.tolist()


def test_scale_log_weights_closed_form():
    rng = np.random.default_rng(2)
    w = np.log(np.abs(rng.standard_cauchy(20)))
    r, s = ol.scale_log_weights(w[:1])
    assert r == w[0] and s[0] == w[0]
    r, s = ol.scale_log_weights(w)
    assert r == pytest.approx(np.log(np.mean(np.exp(w))), rel=1e-14)
    np.testing.assert_allclose(s, np.exp(w - w.max()), rtol=1e-15)
    r, s = ol.scale_log_weights(np.append(w, np.nan))
    assert r == pytest.approx(np.log(np.mean(np.append(np.exp(w), 0))), rel=1e-14)
    assert s[-1] == 0
    r, s = ol.scale_log_weights([-np.inf, -np.inf])
    assert r == -np.inf and list(s) == [-np.inf, -np.inf]
