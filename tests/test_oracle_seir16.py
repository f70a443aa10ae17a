"""The oracle's age-structured SEIR16 pinned without the oracle: its
deterministic mode (binomial -> n p, Exp(rate) -> 1 / rate) against an
independent numpy recursion of the model, its compare_data against scipy's
Poisson log-density, and exact invariants of its stochastic mode.

The model has no counterpart in the reference: force of infection
lambda_a = beta * sum_b C[a, b] I_b / N_b, binomial S->E, E->I, I->R per
age group, incidence reset every 7 days.  A non-symmetric contact matrix
and unequal age-group sizes make a transposed C, or N indexed by the wrong
age, change the result."""
import numpy as np
import pytest
from scipy import stats

import oracle_lib as ol

A = 16


def seir_row(beta, sigma, gamma, I0, N, C, exp_noise=1e6):
    return np.concatenate([[beta, sigma, gamma, exp_noise, I0], N, np.ravel(C)])


def numpy_seir16(row, times, dt, time_start=0.0):
    """Deterministic SEIR16 written from the model definition; state at
    each time (n_times, 80)."""
    beta, sigma, gamma, noise, I0 = row[:5]
    N = row[5:5 + A]
    C = row[5 + A:].reshape(A, A)
    S, E, I, R, inc = N - I0, np.zeros(A), np.full(A, I0), np.zeros(A), np.zeros(A)
    out = []
    t = time_start
    for target in times:
        n = int(round(max(0.0, target - t) / dt))
        for s in range(n):
            if np.fmod(t + s * dt, 7.0) == 0:
                inc = np.zeros(A)
            foi = C @ (I / N)
            n_SE = S * (1 - np.exp(-beta * foi * dt))
            n_EI = E * (1 - np.exp(-sigma * dt))
            n_IR = I * (1 - np.exp(-gamma * dt))
            S, E, I, R, inc = S - n_SE, E + n_SE - n_EI, I + n_EI - n_IR, R + n_IR, inc + n_SE
        t = t + n * dt
        out.append(np.concatenate([S, E, I, R, inc]))
    return np.array(out)


def groups(seed=3, G=3):
    rng = np.random.default_rng(seed)
    rows = []
    for _ in range(G):
        C = rng.uniform(0, 2, (A, A))          # not symmetrised
        N = np.round(rng.uniform(1e3, 1e5, A))
        rows.append(seir_row(rng.uniform(0.2, 0.8), rng.uniform(0.1, 0.5),
                             rng.uniform(0.05, 0.3), float(rng.integers(1, 20)), N, C))
    return np.array(rows)


def test_deterministic_oracle_equals_numpy_recursion():
    rows = groups()
    times = np.arange(7.0, 7.0 * 13, 7.0)
    out, _ = ol.sir_simulate(rows, 2, 0.0, 0.25, times, ol.prng(rows.shape[0] * 2, seed=1),
                             deterministic=True, model="seir16")
    for g in range(rows.shape[0]):
        expect = numpy_seir16(rows[g], times, 0.25)
        np.testing.assert_allclose(out[:, g, 0], expect, rtol=1e-11, atol=1e-9)
        np.testing.assert_array_equal(out[:, g, 0], out[:, g, 1])


def test_recursion_tells_contact_orientation_apart():
    # the check above would fail for a model that used C transposed or N of
    # the infected person's age instead of the infectious one's
    row = groups(seed=4, G=1)[0]
    times = np.arange(7.0, 7.0 * 5, 7.0)
    out, _ = ol.sir_simulate(row[None], 1, 0.0, 0.25, times, ol.prng(1, seed=1),
                             deterministic=True, model="seir16")
    C = row[5 + A:].reshape(A, A)
    wrong = row.copy()
    wrong[5 + A:] = C.T.ravel()
    rel = np.abs(numpy_seir16(wrong, times, 0.25) - out[:, 0, 0]) / (np.abs(out[:, 0, 0]) + 1)
    assert rel.max() > 1e-3


def test_deterministic_likelihood_equals_poisson_sum():
    rows = groups(seed=5, G=2)
    times = np.arange(7.0, 7.0 * 11, 7.0)
    rng = np.random.default_rng(2)
    data = rng.poisson(40, (times.size, 2, A)).astype(float)
    data[3, 0, 5] = np.nan
    data[7, 1, :] = np.nan
    res = ol.discrete_unfilter(rows, 3, 0.0, times, 0.25, data, model="seir16",
                               trajectories=True)
    for g in range(2):
        traj = numpy_seir16(rows[g], times, 0.25)
        lam = traj[:, 64:80] + 1 / rows[g, 3]
        d = data[:, g]
        keep = ~np.isnan(d)
        expect = stats.poisson.logpmf(d[keep], lam[keep]).sum()
        for p in range(3):
            assert res["ll"][g * 3 + p] == pytest.approx(expect, rel=1e-12)
        np.testing.assert_allclose(res["traj"][:, g, 0], traj, rtol=1e-11, atol=1e-9)


def test_stochastic_oracle_conserves_population_and_sums_incidence():
    rows = groups(seed=6, G=2)
    times = np.arange(7.0, 7.0 * 9, 7.0)
    P = 20
    out, _ = ol.sir_simulate(rows, P, 0.0, 0.25, times, ol.prng(2 * P, seed=8),
                             model="seir16")
    for g in range(2):
        N = rows[g, 5:5 + A]
        I0 = rows[g, 4]
        st = out[:, g]                          # (T, P, 80)
        np.testing.assert_array_equal(st[..., :16] + st[..., 16:32] + st[..., 32:48]
                                      + st[..., 48:64], np.broadcast_to(N, st[..., :16].shape))
        # weekly incidence (reset at each observation) sums to the infections
        np.testing.assert_array_equal(st[..., 64:80].sum(axis=0), (N - I0) - st[-1, :, :16])
        assert np.all(st == np.round(st))
