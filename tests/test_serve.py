"""Scheduling policy of the serving loop, on the CPU with stand-in engines:
admission by footprint within the budget, arrival order, requests joining
while others run (continuous batching), every step served exactly once."""
import math

import pytest

import serve


class Clock:
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        return self.t


class FakeEngine:
    def __init__(self, log, rid):
        self.log, self.rid, self.closed = log, rid, False

    def close(self):
        self.closed = True
        self.log.append(("close", self.rid))


class FakeSim:
    def __init__(self, r, log):
        self.r, self.engine, self.steps = r, FakeEngine(log, r.rid), 0

    def schedule(self, n, cluster=False):
        self.steps += n

        class S:
            n_it = n
        return S()


def harness(budget, chunk, step_cost=0.01):
    clock, log, sims = Clock(), [], {}

    def start(r, luts):
        sims[r.rid] = FakeSim(r, log)
        log.append(("start", r.rid, clock.t))
        return sims[r.rid]

    class Batch:
        def __init__(self, engines):
            self.members = [e.rid for e in engines]
            assert not any(e.closed for e in engines)
            log.append(("batch", tuple(self.members)))

        def run(self, scheds):
            clock.t += step_cost * max(s.n_it for s in scheds)

        def close(self):
            pass

    srv = serve.Server(budget, chunk=chunk, start_fn=start, batch_fn=Batch,
                       footprint_fn=lambda r, luts: r.L * 1000, clock=clock)
    return srv, log, sims


def req(rid, L, sweeps, arrival):
    return serve.Request(rid=rid, seed=rid, L=L, Lz=L, bc=0, nt=16,
                         sweeps=sweeps, arrival=arrival)


def test_budget_and_order():
    srv, log, sims = harness(budget=40000, chunk=10)
    reqs = [req(0, 16, 50, 0.0), req(1, 16, 50, 0.0), req(2, 16, 30, 0.0),
            req(3, 8, 20, 0.0)]
    res = srv.run(reqs)
    assert [r.rid for r in res] == [0, 1, 2, 3]
    assert all(sims[r.rid].steps == r.sweeps for r in reqs)
    assert srv.stats["max_bytes"] <= 40000
    # 16000 + 16000 fit, the third waits for memory; arrival order holds
    starts = [e[1] for e in log if e[0] == "start"]
    assert starts == [0, 1, 2, 3]
    assert srv.stats["deferred"] > 0
    assert all(x.finished >= x.admitted >= x.arrival for x in res)


def test_late_arrivals_join_running_batch():
    srv, log, sims = harness(budget=10 ** 9, chunk=5)
    reqs = [req(0, 8, 100, 0.0), req(1, 16, 20, 0.12), req(2, 24, 10, 0.3)]
    res = srv.run(reqs)
    batches = [e[1] for e in log if e[0] == "batch"]
    # the late requests run beside request 0 before it completes
    assert any(set(b) >= {0, 1} for b in batches)
    assert any(set(b) >= {0, 2} for b in batches)
    fin = {x.rid: x.finished for x in res}
    assert fin[1] < fin[0] and fin[2] < fin[0]
    assert all(sims[r.rid].steps == r.sweeps for r in reqs)


def test_oversized_request_is_refused():
    srv, _, _ = harness(budget=5000, chunk=4)
    with pytest.raises(ValueError):
        srv.run([req(0, 16, 10, 0.0)])


def test_oversized_request_is_refused_before_anything_starts():
    srv, log, _ = harness(budget=20000, chunk=4)
    with pytest.raises(ValueError):
        srv.run([req(0, 8, 10, 0.0), req(1, 16, 10, 0.5), req(2, 24, 10, 1.0)])
    assert log == []  # request 2 (24000 bytes) was refused up front


def test_failure_releases_running_engines():
    """A start that fails while other requests run: the error propagates,
    and the running requests' engines and the batch are closed."""
    srv, log, sims = harness(budget=10 ** 9, chunk=5)
    closed_batches = []
    inner_start = srv.start_fn

    def start(r, luts):
        if r.rid == 2:
            raise RuntimeError("device start failed")
        return inner_start(r, luts)

    class Batch:
        def __init__(self, engines):
            self.n = len(engines)

        def run(self, scheds):
            srv.clock.t += 0.01 * max(s.n_it for s in scheds)

        def close(self):
            closed_batches.append(self.n)

    srv.start_fn, srv.batch_fn = start, Batch
    reqs = [req(0, 8, 100, 0.0), req(1, 16, 100, 0.0), req(2, 8, 10, 0.05)]
    with pytest.raises(RuntimeError):
        srv.run(reqs)
    assert sims[0].engine.closed and sims[1].engine.closed
    assert closed_batches == [2]


def test_poisson_requests_shapes():
    rs = serve.poisson_requests(300, rate=50.0, seed=3)
    assert all(r.L in (8, 16, 24, 32) and r.Lz % 8 == 0 and
               r.L <= r.Lz <= 16 * r.L for r in rs)
    assert all(1000 <= r.sweeps <= 100000 for r in rs)
    assert all(b.arrival > a.arrival for a, b in zip(rs, rs[1:]))
    assert {r.bc for r in rs} == {0, 1} and {r.nt for r in rs} == {16, 24, 32}
    mean_gap = rs[-1].arrival / len(rs)
    assert abs(mean_gap - 1 / 50.0) < 0.3 / 50.0
    lg = [math.log(r.sweeps) for r in rs]
    assert abs(sum(lg) / len(lg) - (math.log(1000) + math.log(100000)) / 2) < 0.4


def test_shape_signature_and_cache(tmp_path):
    import eatube

    class E:
        def __init__(self, *k):
            self.L, self.Lz, self.nr, self.nt, self.ns, self.bc, self.ctm = k

    a, b = E(8, 16, 4, 6, 2, 0, 3), E(16, 64, 4, 4, 1, 1, 2)
    assert eatube.shape_signature([a, b]) == eatube.shape_signature([b, a])
    assert eatube.shape_signature([a, b]) != eatube.shape_signature([a, b],
                                                                    ptfr=2)
    assert eatube.shape_signature([a]) != eatube.shape_signature([a],
                                                                 cluster=False)
    path = str(tmp_path / "s.json")
    assert eatube.load_shape_cache(path) == {}
    eatube.save_shape_cache(path, {"k": [32, 512]})
    assert eatube.load_shape_cache(path) == {"k": [32, 512]}
    # a cached mix is answered without building anything
    cache = {eatube.shape_signature([a, b]): [32, 512, 0, 1]}
    assert eatube.tune_shape([b, a], cache=cache) == ((32, 512, 0, 1), {})
    assert eatube.DEFAULT_SHAPE in eatube.SHAPES
    assert len(set(eatube.SHAPES)) == len(eatube.SHAPES) == 36


def test_worker_preparation_serves_every_request_once():
    """With host preparation on worker threads, requests are admitted only
    once prepared, every step is served exactly once, and each start gets
    the request's own preparation."""
    clock, log, sims, prepared = Clock(), [], {}, {}

    def prepare(r, luts):
        prepared[r.rid] = ("prep", r.rid)
        return prepared[r.rid]

    def start(r, luts, prep):
        assert prep == ("prep", r.rid)
        sims[r.rid] = FakeSim(r, log)
        return sims[r.rid]

    class Batch:
        def __init__(self, engines):
            log.append(("batch", tuple(e.rid for e in engines)))

        def run(self, scheds):
            clock.t += 0.01 * max(s.n_it for s in scheds)

        def close(self):
            pass

    srv = serve.Server(10 ** 9, chunk=7, start_fn=start, batch_fn=Batch,
                       footprint_fn=lambda r, luts: 1, clock=clock, workers=3,
                       prepare_fn=prepare)
    reqs = [req(i, 8 * (1 + i % 3), 10 + 7 * i, 0.05 * i) for i in range(9)]
    res = srv.run(reqs)
    assert [r.rid for r in res] == list(range(9))
    assert all(sims[r.rid].steps == r.sweeps for r in reqs)
    assert set(prepared) == set(range(9))
    assert all(x.finished >= x.admitted >= x.arrival for x in res)
