"""Trajectory export following the ancestry on the GPU: the kernel on given
states against forward-built lineages (the reference helper's ground truth),
the filter's export against the oracle's walk over the same saved states,
SIR lineage invariants, and select_random_particle (tests/testthat/test-filter.R:
55-80,120-150,540-570)."""
import numpy as np
import pytest

import oracle_lib as ol
from test_oracle_trajectories import example_trajectories

pytestmark = pytest.mark.gpu

NAMES = ["N", "I0", "beta", "gamma", "exp_noise"]


def test_kernel_matches_forward_tracking(gpu_lib):
    for G in [1, 3]:
        raw, anc, true = example_trajectories(S=6, G=G)
        np.testing.assert_array_equal(gpu_lib.gpu_trajectory_lineage(raw, anc), true)
    raw, anc, true = example_trajectories(S=6, G=3)
    sel = np.array([4, 0, 6])
    out = gpu_lib.gpu_trajectory_lineage(raw, anc, sel)
    for g in range(3):
        np.testing.assert_array_equal(out[:, g, 0], true[:, g, sel[g]])


def filter_with_trajectories(lib, G=3, P=200, n_time=30, precision="fp64", engine="auto",
                             seed=42):
    pars = np.array([[1e4, 10, 0.3 + 0.05 * g, 0.1, 1e6] for g in range(G)])
    time = np.arange(1.0, n_time + 1)
    rng = np.random.default_rng(seed)
    data = rng.poisson(20, (n_time, G)).astype(float)
    data[min(5, n_time - 2)] = np.nan      # a time without resampling
    obj = lib.dust_filter_create(lib.sir(), 0.0, {"time": time, "incidence": data}, P,
                                 n_groups=G, dt=0.25, seed=seed, precision=precision)
    obj.set_engine(engine)
    plist = [dict(zip(NAMES, p)) for p in pars]
    lib.dust_likelihood_run(obj, plist, save_trajectories=True)
    return obj, plist


@pytest.mark.parametrize("engine", ["persistent", "pipeline"])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_filter_export_follows_ancestry(gpu_lib, engine, precision):
    obj, _ = filter_with_trajectories(gpu_lib, precision=precision, engine=engine)
    h = gpu_lib.dust_likelihood_last_trajectories(obj)
    raw = gpu_lib.dust_likelihood_last_trajectories_raw(obj)
    anc = gpu_lib.dust_likelihood_last_ancestry(obj)
    np.testing.assert_array_equal(h, ol.trajectory_lineage(raw, anc))
    np.testing.assert_array_equal(h[-1], gpu_lib.dust_likelihood_last_state(obj))
    # along a lineage the SIR flows only one way; across raw particle slots
    # (reshuffled by resampling) they do not
    S, R, cum = h[..., 0], h[..., 2], h[..., 3]
    assert np.all(np.diff(S, axis=0) <= 0)
    assert np.all(np.diff(R, axis=0) >= 0)
    assert np.all(np.diff(cum, axis=0) >= 0)
    assert np.any(np.diff(raw[..., 3], axis=0) < 0)
    np.testing.assert_array_equal(h[..., 0] + h[..., 1] + h[..., 2], 1e4)


def test_select_random_particle(gpu_lib):
    G, P = 3, 100
    obj, _ = filter_with_trajectories(gpu_lib, G=G, P=P, n_time=10)
    s1 = gpu_lib.dust_likelihood_rng_state(obj).reshape(G, P + 2, 4)
    h1 = gpu_lib.dust_likelihood_last_trajectories(obj)
    h2 = gpu_lib.dust_likelihood_last_trajectories(obj, select_random_particle=True)
    assert h2.shape == (10, G, 1, 5)
    s2 = gpu_lib.dust_likelihood_rng_state(obj).reshape(G, P + 2, 4)
    # one draw from filter stream 2g+1 per group; every other stream untouched
    for g in range(G):
        _, after = ol.rng_next(s1[g, 1], 1)
        assert np.array_equal(s2[g, 1], after)
    assert np.array_equal(np.delete(s2, 1, axis=1), np.delete(s1, 1, axis=1))
    # the chosen particle is floor(u * P) of that draw, and its lineage is
    # the exported lineage of that particle
    for g in range(G):
        u, _ = ol.random_real(s1[g, 1], 1)
        k = int(np.floor(u[0] * P))
        np.testing.assert_array_equal(h2[:, g, 0], h1[:, g, k])
    # cached until the next run
    h3 = gpu_lib.dust_likelihood_last_trajectories(obj, select_random_particle=True)
    np.testing.assert_array_equal(h3, h2)
    assert np.array_equal(gpu_lib.dust_likelihood_rng_state(obj).reshape(G, P + 2, 4), s2)


def test_trajectories_not_current(gpu_lib):
    obj, plist = filter_with_trajectories(gpu_lib, G=1, P=50, n_time=5)
    gpu_lib.dust_likelihood_run(obj, None)
    with pytest.raises(gpu_lib.DustError, match="Trajectories are not current"):
        gpu_lib.dust_likelihood_last_trajectories(obj)


def test_index_group_exports_selected_groups(gpu_lib):
    obj, plist = filter_with_trajectories(gpu_lib, G=4, P=64, n_time=12)
    full = gpu_lib.dust_likelihood_last_trajectories(obj)
    ll = gpu_lib.dust_likelihood_run(obj, None, save_trajectories=True, index_group=[2, 0])
    sub = gpu_lib.dust_likelihood_last_trajectories(obj)
    assert sub.shape == (12, 2, 64, 5)
    assert np.all(np.isfinite(ll))
    assert np.all(sub[..., 0] + sub[..., 1] + sub[..., 2] == 1e4)
    assert full.shape == (12, 4, 64, 5)


def test_snapshots_follow_ancestry(gpu_lib):
    G, P, T = 3, 150, 20
    obj, plist = filter_with_trajectories(gpu_lib, G=G, P=P, n_time=T)
    snaps = np.array([3.0, 10.0, 17.0])
    gpu_lib.dust_likelihood_run(obj, None, save_trajectories=True, save_snapshots=snaps)
    h = gpu_lib.dust_likelihood_last_trajectories(obj)
    s = gpu_lib.dust_likelihood_last_snapshots(obj)
    assert s.shape == (3, G, P, 5)
    np.testing.assert_array_equal(s, h[snaps.astype(int) - 1])
    s1 = gpu_lib.dust_likelihood_last_snapshots(obj, select_random_particle=True)
    h1 = gpu_lib.dust_likelihood_last_trajectories(obj, select_random_particle=True)
    np.testing.assert_array_equal(s1, h1[snaps.astype(int) - 1])
    # the chosen particle is shared with last_state(select_random_particle)
    st = gpu_lib.dust_likelihood_last_state(obj, select_random_particle=True)
    np.testing.assert_array_equal(st, h1[-1, :, 0])
    # snapshots alone: no trajectories, and the next run forgets them
    gpu_lib.dust_likelihood_run(obj, None, save_snapshots=[5.0])
    with pytest.raises(gpu_lib.DustError, match="Trajectories are not current"):
        gpu_lib.dust_likelihood_last_trajectories(obj)
    assert gpu_lib.dust_likelihood_last_snapshots(obj).shape == (1, G, P, 5)
    gpu_lib.dust_likelihood_run(obj, None)
    with pytest.raises(gpu_lib.DustError, match="Snapshots are not current"):
        gpu_lib.dust_likelihood_last_snapshots(obj)
    with pytest.raises(gpu_lib.DustError, match="must be found in 'time'"):
        gpu_lib.dust_likelihood_run(obj, None, save_snapshots=[2.5])
