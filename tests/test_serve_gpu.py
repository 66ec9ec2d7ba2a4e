"""Continuous batching on the GPU: requests that arrive while others run
join them, and every request's final state equals the same request run
alone (bit-exact: spins, PT permutations, energy sums, energies)."""
import sys

import numpy as np
import pytest

import eatube
import serve

pytestmark = pytest.mark.gpu


def alone(r):
    betas = eatube.betas_linear(r.tmin, r.tmax, r.nt)
    sim = eatube.Simulation(r.L, r.Lz, r.nr, r.nt, r.ns, bc=r.bc, betas=betas,
                            seeds=r.seeds(), job_id=r.rid % (1 << 14))
    sim.steps(r.sweeps, cluster=r.cluster)
    es, _ = sim.engine.read_measures()
    out = dict(spins=sim.spins(), perms=sim.engine.download_perms(), esum=es,
               energies=sim.engine.energy())
    sim.engine.close()
    return out


def test_mix_equals_alone():
    reqs = [
        serve.Request(rid=0, seed=11, L=8, Lz=64, bc=1, nt=16, sweeps=90),
        serve.Request(rid=1, seed=12, L=16, Lz=32, bc=0, nt=24, sweeps=40,
                      arrival=0.0),
        serve.Request(rid=2, seed=13, L=24, Lz=48, bc=1, nt=16, sweeps=25,
                      arrival=0.05, cluster=True),
        serve.Request(rid=3, seed=14, L=32, Lz=32, bc=0, nt=32, sweeps=30,
                      arrival=0.1),
        serve.Request(rid=4, seed=15, L=8, Lz=8, bc=0, nt=16, sweeps=70,
                      arrival=0.15, cluster=True),
    ]
    srv = serve.Server(budget=2 << 30, chunk=7, keep=True)
    res = srv.run(reqs)
    print("stats", srv.stats, flush=True)
    assert srv.stats["batches_built"] >= 2
    for r, x in zip(reqs, res):
        print("request", r, "latency %.3f s" % x.latency, flush=True)
        ref = alone(r)
        for k in ("spins", "perms", "esum", "energies"):
            assert np.array_equal(getattr(x, k), ref[k]), (r.rid, k)


def test_poisson_stream_at_the_serving_shapes():
    """Requests drawn as the serving workload draws them (L in {8, 16, 24,
    32}, Lz a multiple of 8 in [L, 16L], PBC or OBC, NT in {16, 24, 32},
    sweeps log-uniform in [1e3, 1e5], Houdayer moves), with sweeps capped at
    1500 to fit the test's time, plus the largest shape (L = 32, Lz = 512,
    NT = 32): each result of the mix equals the request run alone."""
    reqs = serve.poisson_requests(12, rate=40.0, seed=11)
    for r in reqs:
        r.sweeps = min(r.sweeps, 1500)
    reqs.append(serve.Request(rid=12, seed=99, L=32, Lz=512, bc=1, nt=32,
                              sweeps=1000, arrival=reqs[-1].arrival))
    srv = serve.Server(budget=8 << 30, chunk=64, keep=True, workers=4)
    res = srv.run(reqs)
    print("stats", srv.stats, flush=True)
    assert srv.stats["max_running"] >= 2
    for r, x in zip(reqs, res):
        print("request", r.rid, (r.L, r.Lz, r.bc, r.nt, r.sweeps),
              "latency %.3f s" % x.latency, flush=True)
        ref = alone(r)
        for k in ("spins", "perms", "esum", "energies"):
            assert np.array_equal(getattr(x, k), ref[k]), (r.rid, k)
