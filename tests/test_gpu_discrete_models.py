"""One set of discrete-time kernels for every model (models.hpp): the SEIR16
and random-walk models run through the same pipeline, system and compare
kernels as the SIR, with no kernel of their own.

Bars: FP64 draw-for-draw with the oracle (states and streams exact,
log-likelihoods within 1e-12); the FP32 SEIR16, whose force of infection is
a matrix-core contraction, against the FP64 oracle in deterministic mode
within FP32 tolerance (relative 2e-5 over 8 weeks; the bound is stated
with each check), and one force-of-infection evaluation against the naive
FP64 product."""
import numpy as np
import pytest

import oracle_lib as ol

pytestmark = pytest.mark.gpu

A = 16


def seir_rows(G, seed=3, symmetric=False):
    rng = np.random.default_rng(seed)
    rows = []
    for _ in range(G):
        C = rng.uniform(0, 2, (A, A))
        if symmetric:
            C = (C + C.T) / 2
        N = np.round(rng.uniform(2e3, 5e4, A))
        rows.append(np.concatenate([[rng.uniform(0.3, 0.9), rng.uniform(0.15, 0.4),
                                     rng.uniform(0.1, 0.3), 1e6, float(rng.integers(2, 12))],
                                    N, C.ravel()]))
    return np.array(rows)


def as_pars(row):
    return {"beta": row[0], "sigma": row[1], "gamma": row[2], "exp_noise": row[3], "I0": row[4],
            "N": row[5:5 + A], "C": row[5 + A:].reshape(A, A)}


def seir_data(row, times, seed=9):
    sim, _ = ol.sir_simulate(row[None, :], 1, 0.0, 0.25, times, ol.prng(1, seed=seed),
                             model="seir16")
    return sim[:, 0, 0, 64:80].copy()


# ---- SEIR16 -------------------------------------------------------------------
def test_seir16_fp64_filter_draw_for_draw(gpu_lib):
    G, P = 2, 96
    rows = seir_rows(G)
    times = np.arange(7.0, 7.0 * 7, 7.0)
    data = np.stack([seir_data(rows[g], times, seed=20 + g) for g in range(G)], axis=1)
    data[2, 1, 3] = np.nan
    obj = gpu_lib.dust_filter_create(gpu_lib.seir16(), 0.0, {"time": times, "incidence": data},
                                     P, n_groups=G, dt=0.25, seed=17, precision="fp64")
    ll = gpu_lib.dust_likelihood_run(obj, [as_pars(r) for r in rows], save_trajectories=True)
    ref = ol.sir_filter(rows, P, 0.0, times, 0.25, data, seed=17, ancestry=True, n_threads=8,
                        model="seir16")
    np.testing.assert_allclose(ll, ref["ll"], rtol=1e-12)
    assert np.array_equal(gpu_lib.dust_likelihood_last_state(obj), ref["state"])
    assert np.array_equal(gpu_lib.dust_likelihood_last_ancestry(obj), ref["ancestry"])
    assert np.array_equal(gpu_lib.dust_likelihood_rng_state(obj).reshape(-1, 4), ref["rng"])


def test_seir16_fp64_system_draw_for_draw(gpu_lib):
    G, P = 2, 70
    rows = seir_rows(G, seed=5)
    sys = gpu_lib.dust_system_create(gpu_lib.seir16(), [as_pars(r) for r in rows], P,
                                     n_groups=G, dt=0.25, seed=4, precision="fp64")
    gpu_lib.dust_system_set_state_initial(sys)
    times = np.arange(0.0, 43.0, 7.0)
    got = gpu_lib.dust_system_simulate(sys, times)
    expect, rng_after = ol.sir_simulate(rows, P, 0.0, 0.25, times, ol.prng(G * P, seed=4),
                                        model="seir16")
    assert np.array_equal(got, expect)
    assert np.array_equal(gpu_lib.dust_system_rng_state(sys).reshape(-1, 4), rng_after)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_seir16_deterministic_system_matches_oracle(gpu_lib, precision):
    # FP64: the CPU's arithmetic order, so within rounding of the oracle;
    # FP32: state, samplers and the matrix-core force of infection in FP32,
    # 8 weeks (224 steps) of a growing epidemic: relative 2e-5 in every
    # state and incidence (the counts reach ~1e4, so FP32 resolution alone
    # is ~1e-7 per step)
    G, P = 3, 130
    rows = seir_rows(G, seed=11)
    sys = gpu_lib.dust_system_create(gpu_lib.seir16(), [as_pars(r) for r in rows], P,
                                     n_groups=G, dt=0.25, seed=1, deterministic=True,
                                     precision=precision)
    gpu_lib.dust_system_set_state_initial(sys)
    times = np.arange(7.0, 7.0 * 9, 7.0)
    got = gpu_lib.dust_system_simulate(sys, times)
    expect, _ = ol.sir_simulate(rows, 1, 0.0, 0.25, times, ol.prng(G, seed=1),
                                deterministic=True, model="seir16")
    expect = np.broadcast_to(expect, got.shape)
    if precision == "fp64":
        np.testing.assert_allclose(got, expect, rtol=1e-12, atol=1e-9)
    else:
        np.testing.assert_allclose(got, expect, rtol=2e-5, atol=1e-3)


def test_seir16_fp32_deterministic_likelihood_matches_oracle(gpu_lib):
    # the deterministic FP32 model through compare_data: every particle of
    # every group scores the oracle's deterministic log-density within a
    # relative 2e-5 (a sum over 16 ages of Poisson terms of FP32 incidence)
    G, P = 2, 64
    rows = seir_rows(G, seed=13)
    times = np.arange(7.0, 7.0 * 9, 7.0)
    sys = gpu_lib.dust_system_create(gpu_lib.seir16(), [as_pars(r) for r in rows], P,
                                     n_groups=G, dt=0.25, seed=1, deterministic=True,
                                     precision="fp32")
    gpu_lib.dust_system_set_state_initial(sys)
    data = np.random.default_rng(3).poisson(50, (times.size, G, A)).astype(float)
    ref = ol.discrete_unfilter(rows, 1, 0.0, times, 0.25, data, model="seir16")
    ll = np.zeros((G, P))
    for i, t in enumerate(times):
        gpu_lib.dust_system_run_to_time(sys, t)
        ll += gpu_lib.dust_system_compare_data(sys, data[i])
    np.testing.assert_allclose(ll, np.broadcast_to(ref["ll"][:, None], ll.shape), rtol=2e-5)


def test_seir16_force_of_infection_on_matrix_cores_equals_naive_product(gpu_lib):
    # one deterministic FP32 step from states with a different infectious
    # count in every age of every particle: the incidence of the step is
    # S_a (1 - exp(-lambda_a dt)), so lambda_a of each particle is recovered
    # and compared with beta * sum_b C[a, b] I_b / N_b in FP64.  Particles
    # span whole wavefronts (64) and a partial one, so every lane position of
    # the MFMA operands is exercised.
    G, P = 2, 150
    rows = seir_rows(G, seed=21)
    rng = np.random.default_rng(4)
    state = np.zeros((G, P, 80))
    for g in range(G):
        N = rows[g, 5:5 + A]
        I = rng.uniform(0, 0.05, (P, A)) * N
        state[g, :, 32:48] = np.round(I)
        state[g, :, 0:16] = N - state[g, :, 32:48]
    dt = 0.25
    sys = gpu_lib.dust_system_create(gpu_lib.seir16(), [as_pars(r) for r in rows], P,
                                     n_groups=G, dt=dt, time=1.0, seed=1, deterministic=True,
                                     precision="fp32")
    gpu_lib.dust_system_set_state(sys, state)
    gpu_lib.dust_system_run_to_time(sys, 1.0 + dt)
    after = gpu_lib.dust_system_state(sys)
    for g in range(G):
        beta, N, C = rows[g, 0], rows[g, 5:5 + A], rows[g, 5 + A:].reshape(A, A)
        S = state[g, :, 0:16]
        n_SE = after[g, :, 64:80]
        lam_gpu = -np.log1p(-n_SE / S) / dt
        lam = beta * (state[g, :, 32:48] / N) @ C.T      # lam[p, a] = sum_b C[a, b] x[p, b]
        np.testing.assert_allclose(lam_gpu, lam, rtol=2e-5)
        # the transposed contraction is far from it
        lam_t = beta * (state[g, :, 32:48] / N) @ C
        assert np.max(np.abs(lam_t - lam_gpu) / lam) > 1e-2


# ---- random walk (inst/examples/walk.cpp) ------------------------------------
@pytest.mark.parametrize("precision", ["fp64"])
def test_walk_system_draw_for_draw(gpu_lib, precision):
    # test-walk.R:1-22,195-212: random initial state and several walkers
    G, P = 3, 40
    pars = [{"sd": float(g + 1), "len": 3, "random_initial": 1.0} for g in range(G)]
    sys = gpu_lib.dust_system_create(gpu_lib.walk(), pars, P, n_groups=G, dt=0.5, seed=42,
                                     precision=precision)
    gpu_lib.dust_system_set_state_initial(sys)
    rows = np.array([[p["sd"], 3.0, 1.0] for p in pars])
    times = np.array([0.0, 1.0, 2.5, 5.0])
    got = gpu_lib.dust_system_simulate(sys, times)
    rng0 = ol.prng(G * P, seed=42)
    expect, rng_after = ol.walk_simulate(rows, P, 0.0, 0.5, times, rng0)
    np.testing.assert_allclose(got, expect, rtol=1e-14, atol=1e-14)
    assert np.array_equal(gpu_lib.dust_system_rng_state(sys).reshape(-1, 4), rng_after)
    assert got.shape == (4, G, P, 3)


def test_walk_deterministic_and_fixed_initial(gpu_lib):
    sys = gpu_lib.dust_system_create(gpu_lib.walk(), {"sd": 1.0}, 10, seed=1,
                                     deterministic=True)
    gpu_lib.dust_system_set_state_initial(sys)
    assert np.all(gpu_lib.dust_system_state(sys) == 0)
    s = np.arange(10.0).reshape(1, 10, 1)
    gpu_lib.dust_system_set_state(sys, s)
    gpu_lib.dust_system_run_to_time(sys, 5)
    np.testing.assert_array_equal(gpu_lib.dust_system_state(sys), s)


def test_walk_statistics_and_update_pars(gpu_lib):
    # test-walk.R:259-296: the variance after t steps of sd * dt is
    # t * (sd dt)^2; sd is updatable, len is not
    P = 20000
    sys = gpu_lib.dust_system_create(gpu_lib.walk(), {"sd": 1.0}, P, dt=0.25, seed=3)
    gpu_lib.dust_system_set_state_initial(sys)
    gpu_lib.dust_system_update_pars(sys, {"sd": 2.0})
    gpu_lib.dust_system_run_to_time(sys, 10)
    x = gpu_lib.dust_system_state(sys)[0, :, 0]
    var = 40 * (2.0 * 0.25) ** 2
    assert abs(x.mean()) < 4 * np.sqrt(var / P)
    assert abs(x.var() / var - 1) < 0.05
    with pytest.raises(ValueError, match="constant"):
        gpu_lib.dust_system_update_pars(sys, {"len": 2.0})


def test_walk_refuses_different_state_lengths_and_filters(gpu_lib):
    pars = [{"sd": 1.0, "len": 2.0}, {"sd": 2.0, "len": 2.0}, {"sd": 3.0, "len": 3.0}]
    with pytest.raises(gpu_lib.DustError, match="State length for group 3 was different"):
        gpu_lib.dust_system_create(gpu_lib.walk(), pars, 10, n_groups=3)
    import ctypes as C
    from dust2_amd import _lib
    row = np.array([1.0, 1.0, 0.0])
    time = np.array([1.0])
    out = C.c_void_p()
    buf = C.create_string_buffer(1024)
    rc = _lib.lib().dust2_filter_discrete_alloc(b"walk", _lib.dptr(row), 1, 10, 0.0,
                                                _lib.dptr(time), 1, 1.0, _lib.dptr(time), 1,
                                                None, 0, 1, 0, C.byref(out), buf, len(buf))
    assert rc != 0
    assert "does not support 'compare_data'" in buf.value.decode()
