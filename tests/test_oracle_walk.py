"""Pins the oracle's random walk, the restatement of inst/examples/walk.cpp
(initial: Normal(0, sd) per walker when random_initial, else 0; update:
x_i <- Normal(x_i, sd * dt)), by the law it must follow: after n steps each
walker's displacement is Normal(0, sd * dt * sqrt(n)) and independent of
its start.  Checked by Kolmogorov-Smirnov tests at the 0.1 % level and by
the walkers' independence (correlation of two walkers of one particle);
the deterministic walk stays at its start."""
import numpy as np
from scipy import stats

import oracle_lib as ol


def test_walk_follows_its_normal_law():
    G, P, L = 2, 20000, 3
    rows = np.array([[1.0, L, 0.0], [2.5, L, 1.0]])          # sd, len, random_initial
    dt = 0.5
    times = np.array([0.0, 10.0])
    out, rng_after = ol.walk_simulate(rows, P, 0.0, dt, times, ol.prng(G * P, seed=1))
    assert out.shape == (2, G, P, L)
    n = int(round(10.0 / dt))
    for g in range(G):
        sd, random_initial = rows[g, 0], rows[g, 2]
        start = out[0, g]
        if random_initial:
            assert stats.kstest(start.ravel() / sd, "norm").pvalue > 1e-3
        else:
            assert np.all(start == 0)
        step = out[1, g] - start
        assert stats.kstest(step.ravel() / (sd * dt * np.sqrt(n)), "norm").pvalue > 1e-3
        # walkers of one particle, and the step and its start, are independent
        assert abs(np.corrcoef(step[:, 0], step[:, 1])[0, 1]) < 4 / np.sqrt(P)
        if random_initial:
            assert abs(np.corrcoef(step[:, 0], start[:, 0])[0, 1]) < 4 / np.sqrt(P)
    # every stream moved (one normal draw per walker per step, plus the start)
    assert not np.array_equal(rng_after, ol.prng(G * P, seed=1))


def test_deterministic_walk_stays_at_its_start():
    rows = np.array([[1.0, 2.0, 1.0]])
    out, _ = ol.walk_simulate(rows, 5, 0.0, 1.0, np.array([0.0, 7.0]), ol.prng(5, seed=2),
                              deterministic=True)
    np.testing.assert_array_equal(out[1], out[0])
