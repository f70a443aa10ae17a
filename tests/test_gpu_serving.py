"""Continuous-batching server: every result equals the same request run
alone, run with other chunk sizes, and run through the plain filter entry
points (same kernels, same seed) -- bit for bit, FP32 filters and FP64 ODE
unfilters, up to 8192 particles and 730 intervals; a chain's requests
continue its streams as consecutive runs of one filter seeded at the
chain's long jump; requests wait for memory rather than fail; a failing
request reports its error without disturbing the others."""
import numpy as np
import pytest

import oracle_lib as ol

pytestmark = pytest.mark.gpu

SIR_NAMES = ["N", "I0", "beta", "gamma", "exp_noise"]


def sir_request(rng, T):
    pars = {"N": float(np.round(rng.uniform(1e3, 1e5))), "I0": 10.0,
            "beta": rng.uniform(0.2, 0.5), "gamma": rng.uniform(0.05, 0.2), "exp_noise": 1e6}
    time = np.arange(1.0, T + 1)
    data = rng.poisson(20, T).astype(float)
    return pars, {"time": time, "incidence": data}


def seir_request(rng, T):
    C = rng.uniform(0, 2, (16, 16))
    C = (C + C.T) / 2
    C /= np.max(np.abs(np.linalg.eigvals(C)))
    pars = {"beta": rng.uniform(0.3, 0.8), "sigma": 0.3, "gamma": 0.2, "exp_noise": 1e6,
            "I0": 5.0, "N": np.full(16, 2e4), "C": C}
    time = np.arange(7.0, 7.0 * (T + 1), 7.0)
    data = rng.poisson(30, (T, 16)).astype(float)
    return pars, {"time": time, "incidence": data}


def ode_request(rng, T):
    pars = {"N": 1000.0, "I0": 10.0, "beta": rng.uniform(0.1, 0.5),
            "gamma": rng.uniform(0.05, 0.2)}
    time = np.arange(1.0, T + 1)
    data = rng.poisson(6, T).astype(float)
    return pars, {"time": time, "incidence": data}


def make_mix(lib, seed=1, n=24):
    rng = np.random.default_rng(seed)
    reqs = []
    for i in range(n):
        kind = i % 3
        if kind == 0:
            p, d = sir_request(rng, int(rng.integers(20, 60)))
            reqs.append((lib.sir(), p, int(rng.choice([128, 512, 2048])), d, 0.25, 100 + i))
        elif kind == 1:
            p, d = seir_request(rng, int(rng.integers(3, 8)))
            reqs.append((lib.seir16(), p, int(rng.choice([128, 512])), d, 0.25, 100 + i))
        else:
            p, d = ode_request(rng, int(rng.integers(30, 90)))
            reqs.append((lib.sirode(), p, 1, d, 1.0, 100 + i))
    return reqs


def serve(lib, reqs, chunk=8, hbm=1 << 30):
    from dust2_amd.serving import Server
    sv = Server(hbm_bytes=hbm, chunk_intervals=chunk)
    ids = [sv.submit(g, p, P, 0.0, d, dt=dt, seed=s) for (g, p, P, d, dt, s) in reqs]
    res = {r[0]: r for r in sv.drain()}
    assert sv.memory()[0] == 0
    return [res[i] for i in ids]


def test_results_identical_alone_in_mix_and_chunked(gpu_lib):
    reqs = make_mix(gpu_lib)
    mixed = serve(gpu_lib, reqs, chunk=8)
    assert all(r[2] == 0 for r in mixed)
    rechunked = serve(gpu_lib, reqs[::-1], chunk=3)[::-1]
    for a, b in zip(mixed, rechunked):
        assert a[1] == b[1]
    for k in [0, 1, 2, 5]:
        alone = serve(gpu_lib, [reqs[k]], chunk=50)
        assert alone[0][1] == mixed[k][1]


def test_uneven_particle_counts_share_one_launch(gpu_lib):
    """Requests of one model run in one advance launch per interval whatever
    their particle counts (a block map lays each request's particles onto
    the launch's workgroups) and data lengths.  A larger request admitted
    while smaller ones are part-way through their series enlarges the
    captured launches (more blocks, a wider resample workgroup) without
    changing anyone's result."""
    from dust2_amd.serving import Server
    rng = np.random.default_rng(5)
    small = []
    for i, P in enumerate([64, 300, 1000, 130]):
        p, d = sir_request(rng, int(rng.integers(30, 50)))
        small.append((gpu_lib.sir(), p, P, d, 0.25, 200 + i))
    p, d = sir_request(rng, 25)
    big = (gpu_lib.sir(), p, 5000, d, 0.25, 300)
    sv = Server(hbm_bytes=1 << 30, chunk_intervals=4)
    ids = [sv.submit(g, pp, P, 0.0, dd, dt=dt, seed=s) for (g, pp, P, dd, dt, s) in small]
    sv.step()
    sv.step()
    ids.append(sv.submit(*big[:3], 0.0, big[3], dt=big[4], seed=big[5]))
    res = {r[0]: r for r in sv.drain()}
    got = [res[i] for i in ids]
    assert all(r[2] == 0 for r in got)
    for req, r in zip(small + [big], got):
        alone = serve(gpu_lib, [req], chunk=7)
        assert alone[0][1] == r[1]


def test_server_matches_plain_entry_points(gpu_lib):
    reqs = make_mix(gpu_lib, seed=2, n=6)
    served = serve(gpu_lib, reqs)
    for (g, p, P, d, dt, s), r in zip(reqs, served):
        if g.name == "sirode":
            obj = gpu_lib.dust_unfilter_create(g, 0.0, d, n_particles=1)
            ll = obj.run(p)
        else:
            obj = gpu_lib.dust_filter_create(g, 0.0, d, P, dt=dt, seed=s, precision="fp32")
            obj.set_engine("pipeline", graph=False)
            ll = gpu_lib.dust_likelihood_run(obj, p)
        assert r[1] == ll, (g.name, r[1], ll)


def test_sir_requests_statistically_match_oracle(gpu_lib):
    rng = np.random.default_rng(5)
    p, d = sir_request(rng, 30)
    reqs = [(gpu_lib.sir(), p, 512, d, 0.25, 1000 + i) for i in range(60)]
    ll = np.array([r[1] for r in serve(gpu_lib, reqs)])
    row = np.array([[p[k] for k in SIR_NAMES]] * 40)
    ref = ol.sir_filter(row, 512, 0.0, d["time"], 0.25, d["incidence"][:, None], seed=3,
                        n_threads=16)["ll"]
    se = np.sqrt(ll.var(ddof=1) / ll.size + ref.var(ddof=1) / ref.size)
    assert abs(ll.mean() - ref.mean()) < 3 * se


def test_admission_waits_for_memory(gpu_lib):
    rng = np.random.default_rng(3)
    reqs = []
    for i in range(12):
        p, d = sir_request(rng, 40)
        reqs.append((gpu_lib.sir(), p, 2048, d, 0.25, i))
    tight = serve(gpu_lib, reqs, chunk=5, hbm=2 << 20)   # room for ~2 requests
    roomy = serve(gpu_lib, reqs, chunk=5, hbm=1 << 30)
    assert [r[1] for r in tight] == [r[1] for r in roomy]


def test_failing_request_is_isolated(gpu_lib):
    from dust2_amd.serving import Server
    rng = np.random.default_rng(4)
    good_p, d = sir_request(rng, 20)
    bad_p = dict(good_p, gamma=-0.5)
    sv = Server(hbm_bytes=1 << 28, chunk_intervals=4)
    a = sv.submit(gpu_lib.sir(), good_p, 128, 0.0, d, dt=0.25, seed=1)
    b = sv.submit(gpu_lib.sir(), bad_p, 128, 0.0, d, dt=0.25, seed=2)
    res = {r[0]: r for r in sv.drain()}
    assert res[a][2] == 0 and np.isfinite(res[a][1])
    assert res[b][2] == 1 and "particles reported errors" in res[b][3]
    alone = serve(gpu_lib, [(gpu_lib.sir(), good_p, 128, d, 0.25, 1)], chunk=4)
    assert alone[0][1] == res[a][1]


def test_rng_jump_and_long_jump_match_oracle(gpu_lib):
    from dust2_amd.interface import rng_jump
    st = ol.prng(5, seed=17)
    for long in (False, True):
        got = rng_jump(st, long_jump=long)
        for i in range(5):
            assert np.array_equal(got[i], ol.jump(st[i], long=long))


def test_chain_streams_continue_across_requests(gpu_lib):
    # a chain's requests are the runs of one filter object whose streams
    # start at long_jump^chain of the seed (R/monty.R:29-36): two requests
    # of chain 3 equal two consecutive runs of the plain filter seeded there
    from dust2_amd.serving import Server
    rng = np.random.default_rng(21)
    p1, d = sir_request(rng, 30)
    p2 = dict(p1, beta=p1["beta"] * 1.1)
    start = ol.seed_words(77).reshape(1, 4)
    for _ in range(3):
        start = ol.jump(start[0], long=True).reshape(1, 4)
    plain = gpu_lib.dust_filter_create(gpu_lib.sir(), 0.0, d, 512, dt=0.25, seed=start,
                                       precision="fp32")
    plain.set_engine("pipeline", graph=False)
    ll1 = gpu_lib.dust_likelihood_run(plain, p1)
    ll2 = gpu_lib.dust_likelihood_run(plain, p2)
    sv = Server(hbm_bytes=1 << 28, chunk_intervals=7)
    a = sv.submit(gpu_lib.sir(), p1, 512, 0.0, d, dt=0.25, seed=77, chain=3)
    b = sv.submit(gpu_lib.sir(), p2, 512, 0.0, d, dt=0.25, seed=77, chain=3)
    other = sv.submit(gpu_lib.sir(), p1, 512, 0.0, d, dt=0.25, seed=77, chain=4)
    res = {r[0]: r for r in sv.drain()}
    assert res[a][1] == ll1 and res[b][1] == ll2
    assert res[other][1] != ll1


@pytest.mark.parametrize("model,P,T", [("sir", 8192, 730), ("seir16", 8192, 104),
                                       ("sir", 128, 730)])
def test_largest_requests_equal_plain_filter(gpu_lib, model, P, T):
    rng = np.random.default_rng(P + T)
    if model == "sir":
        g, (p, d) = gpu_lib.sir(), sir_request(rng, T)
    else:
        g, (p, d) = gpu_lib.seir16(), seir_request(rng, T)
    q, e = sir_request(rng, 60)
    served = serve(gpu_lib, [(g, p, P, d, 0.25, 5), (gpu_lib.sir(), q, 512, e, 0.25, 6)],
                   chunk=32)
    obj = gpu_lib.dust_filter_create(g, 0.0, d, P, dt=0.25, seed=5, precision="fp32")
    obj.set_engine("pipeline", graph=False)
    assert served[0][1] == gpu_lib.dust_likelihood_run(obj, p)
