"""Pins the oracle's trajectory export (the backward walk of
dust2/trajectories.hpp:128-189) against lineages built forwards, the way
the reference's own test helper builds its ground truth
(tests/testthat/helper-trajectories.R:1-40): at each resampling the whole
history of every particle is replaced by its ancestor's history."""
import numpy as np

import oracle_lib as ol


def example_trajectories(T=11, S=3, P=7, G=1, seed=1):
    rng = np.random.default_rng(seed)
    raw = np.zeros((T, G, P, S))
    anc = np.tile(np.arange(P, dtype=np.uint32), (T, G, 1))
    true = np.full((T, G, P, S), np.nan)
    offset = 0
    for i in range(T):
        s = offset + np.arange(G * P * S, dtype=float).reshape(G, P, S) + rng.random((G, P, S))
        offset += G * P * S
        if i > 0 and i % 2 == 1:          # resampling at every other time
            for j in range(G):
                k = rng.integers(0, P, P)
                anc[i, j] = k
                s[j] = s[j, k] + 1
                true[:i, j] = true[:i, j][:, k]
        raw[i] = s
        true[i] = s
    return raw, anc, true


def test_lineage_matches_forward_tracking():
    for G in [1, 3]:
        raw, anc, true = example_trajectories(S=6, G=G)
        out = ol.trajectory_lineage(raw, anc)
        np.testing.assert_array_equal(out, true)


def test_single_particle_lineage_matches_forward_tracking():
    raw, anc, true = example_trajectories(S=6, G=3)
    sel = np.array([4, 0, 6])
    out = ol.trajectory_lineage(raw, anc, sel)
    for g in range(3):
        np.testing.assert_array_equal(out[:, g, 0], true[:, g, sel[g]])


def test_lineage_without_resampling_is_the_raw_states():
    raw, _, _ = example_trajectories(S=2, G=2)
    np.testing.assert_array_equal(ol.trajectory_lineage(raw, None), raw)
    ident = np.tile(np.arange(7, dtype=np.uint32), (11, 2, 1))
    np.testing.assert_array_equal(ol.trajectory_lineage(raw, ident), raw)
