"""Pins the oracle's samplers by their distributions (the draw sequences of
monty are not in the reference tree: parity unpinned at that boundary) and
the Poisson log-density by scipy's closed form."""
import numpy as np
import pytest
from scipy import stats

import oracle_lib as ol


def chisq_pvalue(draws, pmf, support):
    counts = np.array([(draws == k).sum() for k in support], dtype=float)
    expect = pmf(support) * draws.size
    # pool cells with small expectation into the neighbouring tail
    keep = expect >= 5
    obs = np.append(counts[keep], draws.size - counts[keep].sum())
    exp = np.append(expect[keep], draws.size - expect[keep].sum())
    if exp[-1] < 5:
        obs[-2] += obs[-1]
        exp[-2] += exp[-1]
        obs, exp = obs[:-1], exp[:-1]
    return stats.chisquare(obs, exp).pvalue


@pytest.mark.parametrize("n,p", [(20, 0.1), (9, 0.5), (5, 0.93), (1000, 0.3), (60, 0.8),
                                 (1_000_000, 0.001), (200, 0.05)])
def test_binomial_distribution(n, p):
    # n*p < 10 uses inversion, >= 10 BTRS; p > 1/2 is drawn as n - Bin(n, 1-p)
    s = ol.seed_words(1234 + n)
    draws, _ = ol.sample("binomial", s, np.full(40000, float(n)), p)
    assert np.all(draws == np.round(draws))
    assert draws.min() >= 0 and draws.max() <= n
    mean = n * p
    sd = np.sqrt(n * p * (1 - p))
    lo, hi = int(max(0, mean - 8 * sd)), int(min(n, mean + 8 * sd)) + 1
    support = np.arange(lo, hi)
    assert chisq_pvalue(draws, lambda k: stats.binom.pmf(k, n, p), support) > 1e-4
    assert abs(draws.mean() - mean) < 5 * sd / np.sqrt(draws.size)


def test_binomial_special_cases():
    s = ol.seed_words(3)
    out, s2 = ol.sample("binomial", s, [0.0, 10.0, 10.0], [0.4, 0.0, 1.0])
    assert list(out) == [0.0, 0.0, 10.0]
    assert np.array_equal(s, s2)  # no uniform consumed
    out, _ = ol.sample("binomial", s, [10.0], [0.25], deterministic=True)
    assert out[0] == 2.5
    with pytest.raises(ol.OracleError, match="Invalid call to binomial with n = 10, p = -0.1, "
                                             "q = 1.1"):
        ol.sample("binomial", s, [10.0], [-0.1])


def test_exponential_distribution():
    s = ol.seed_words(99)
    draws, _ = ol.sample("exponential", s, np.full(20000, 2.5))
    assert stats.kstest(draws, stats.expon(scale=1 / 2.5).cdf).pvalue > 1e-4
    det, _ = ol.sample("exponential", s, [1e6], deterministic=True)
    assert det[0] == 1e-6


def test_normal_distribution():
    s = ol.seed_words(5)
    draws, _ = ol.sample("normal", s, np.full(20000, 1.0), 2.0)
    assert stats.kstest(draws, stats.norm(1, 2).cdf).pvalue > 1e-4


@pytest.mark.parametrize("lam", [0.5, 3.0, 12.0, 250.0])
def test_poisson_distribution(lam):
    s = ol.seed_words(int(lam * 10))
    draws, _ = ol.sample("poisson", s, np.full(30000, lam))
    support = np.arange(0, int(lam + 10 * np.sqrt(lam) + 10))
    assert chisq_pvalue(draws, lambda k: stats.poisson.pmf(k, lam), support) > 1e-4


def test_dpois_log_matches_scipy():
    for x in [0, 1, 5, 30, 200, 5000]:
        for lam in [1e-6, 0.3, 4.0, 30.0, 1234.5, 1e5]:
            assert ol.dpois_log(x, lam) == pytest.approx(stats.poisson.logpmf(x, lam),
                                                         rel=1e-12, abs=1e-12)
    assert ol.dpois_log(0, 0) == 0
    assert ol.dpois_log(3, 0) == -np.inf
