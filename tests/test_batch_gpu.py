"""The batched form: several tubes of different L, Lz, boundary condition,
temperatures and samples advancing in the same launches (one step captured
in a HIP graph and replayed) end bit-identical to each tube run alone, and
to the oracle."""
import sys

import numpy as np
import pytest

import eatube
import orc

pytestmark = pytest.mark.gpu

SEEDS = lambda k: (1000 + 4 * k, 1001 + 4 * k, 1002 + 4 * k, 1003 + 4 * k)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def make(spec, k):
    L, Lz, bc, nt, ns = spec
    betas = eatube.betas_linear(0.6, 1.6, nt)
    return eatube.Simulation(L, Lz, 4, nt, ns, bc=bc, betas=betas,
                             seeds=SEEDS(k), job_id=k)


def state(sim):
    e = sim.engine
    st, acc, att, raw = e.download_pt_state()
    es, hist = e.read_measures(rows=4)
    return dict(spins=sim.spins(), perms=e.download_perms(), pt=bytes(st),
                acc=acc, att=att, raw=raw, esum=es, hist=hist,
                ener=e.energy())


def same(a, b, what):
    for k in a:
        if isinstance(a[k], bytes):
            assert a[k] == b[k], "%s: %s differ" % (what, k)
        else:
            assert np.array_equal(a[k], b[k]), "%s: %s differ" % (what, k)


MIX = [
    (8, 16, 0, 6, 2),     # small PBC
    (8, 64, 1, 8, 1),     # OBC, Lz = 8L
    (16, 16, 1, 5, 2),    # OBC cube
    (16, 64, 0, 4, 1),    # PBC, Lz = 4L
    (24, 24, 0, 4, 1),    # L = 24
    (24, 192, 1, 4, 1),   # slab-path Houdayer (labels do not fit in LDS)
]


def test_batch_equals_tube_by_tube():
    plain = [make(s, k) for k, s in enumerate(MIX)]
    twins = [make(s, k) for k, s in enumerate(MIX)]
    say("tubes", MIX, "slab planes", [t.engine.ctm for t in twins])
    batch = eatube.Batch([t.engine for t in twins])
    for rnd, n_it in enumerate((3, 2)):
        scheds = [t.schedule(n_it, cluster=True, it0=rnd * 3, frec_trw=2)
                  for t in twins]
        offs = batch.run(scheds)
        for k, p in enumerate(plain):
            sch = p.schedule(n_it, cluster=True, it0=rnd * 3, frec_trw=2)
            p.engine.run(sch)
            assert offs[k] == p.rng_offset, "coin offset of tube %d" % k
        for k, (p, t) in enumerate(zip(plain, twins)):
            say("round", rnd, "tube", k, MIX[k])
            same(state(p), state(t), "round %d tube %d %s" % (rnd, k, MIX[k]))
    batch.close()


def test_batch_uneven_steps_and_no_cluster_tubes():
    """Tubes with different step counts, some without cluster moves."""
    specs = [(8, 16, 0, 6, 2), (16, 32, 1, 6, 1), (8, 32, 1, 4, 3)]
    plain = [make(s, 10 + k) for k, s in enumerate(specs)]
    twins = [make(s, 10 + k) for k, s in enumerate(specs)]
    batch = eatube.Batch([t.engine for t in twins])
    n_its, clus = (4, 1, 3), (True, False, True)
    scheds = [t.schedule(n, cluster=c) for t, n, c in zip(twins, n_its, clus)]
    batch.run(scheds)
    for p, n, c in zip(plain, n_its, clus):
        p.steps(n, cluster=c)
    for k, (p, t) in enumerate(zip(plain, twins)):
        same(state(p), state(t), "tube %d %s" % (k, specs[k]))
    batch.close()


def test_batch_matches_oracle():
    specs = [(8, 16, 1, 6, 2), (16, 16, 0, 4, 1)]
    sims = [make(s, 20 + k) for k, s in enumerate(specs)]
    oracles = []
    for k, (L, Lz, bc, nt, ns) in enumerate(specs):
        oracles.append(orc.Sim(L, Lz, 4, nt, ns, bc=bc,
                               betas=orc.betas_linear(0.6, 1.6, nt),
                               seeds=SEEDS(20 + k), job_id=20 + k))
    batch = eatube.Batch([s.engine for s in sims])
    batch.run([s.schedule(2, cluster=True) for s in sims])
    for o in oracles:
        o.steps(2, cluster=True)
    for k, (s, o) in enumerate(zip(sims, oracles)):
        assert np.array_equal(s.spins(), o.spins()), "tube %d spins" % k
        assert np.array_equal(s.engine.download_perms().ravel(),
                              o.field("wctb")), "tube %d perms" % k
        es, _ = s.engine.read_measures()
        assert np.array_equal(es.ravel(), o.field("esum")), "tube %d esum" % k
    batch.close()


def test_every_launch_shape_gives_the_same_state():
    """Launch shapes change speed only: each shape of the batch and of a
    single engine ends bit-identical to the default shape."""
    specs = [(8, 16, 0, 6, 2), (16, 64, 1, 4, 1), (24, 24, 0, 4, 1)]
    ref = [make(s, 30 + k) for k, s in enumerate(specs)]
    for p in ref:
        p.steps(2, cluster=True)
    want = [state(p) for p in ref]
    for shp in eatube.SHAPES:
        say("batch shape", shp)
        # an 8x16 tube has too few planes for kind 4's slabs and their halos
        tubes = [k for k in range(len(specs)) if shp[2] != 4 or k > 0]
        twins = [make(specs[k], 30 + k) for k in tubes]
        batch = eatube.Batch([t.engine for t in twins])
        batch.set_shape(*shp)
        batch.run([t.schedule(2, cluster=True) for t in twins])
        for k, t in zip(tubes, twins):
            same(want[k], state(t), "batch shape %s tube %d" % (shp, k))
        batch.close()
    for shp in eatube.SHAPES:
        say("engine shape", shp)
        solo = make(specs[1], 31)
        solo.engine.set_shape(*shp)
        solo.steps(1, cluster=True)
        solo.engine.sweep(solo.counters(2))      # leaves partials pending
        other = (64, shp[1], (shp[2] + 1) % 3, 1 - shp[3])
        solo.engine.set_shape(*other)            # across a pending sweep
        solo.engine.pt()
        p = make(specs[1], 31)
        p.steps(1, cluster=True)
        p.engine.sweep(p.counters(2))
        p.engine.pt()
        same(state(p), state(solo), "engine shape %s" % (shp,))


def test_tuner_picks_a_shape_and_caches_it(tmp_path):
    specs = [(8, 16, 0, 6, 2), (16, 32, 1, 4, 1)]
    sims = [make(s, 40 + k) for k, s in enumerate(specs)]
    before = [state(s) for s in sims]
    cache = {}
    best, times = eatube.tune_shape([s.engine for s in sims], steps=2,
                                    cache=cache)
    say("tuned", best, times)
    assert best in eatube.SHAPES and best in times
    assert set(times) <= set(eatube.SHAPES) and len(times) >= 6
    assert min(times.values()) == times[best]
    for b, s in zip(before, sims):     # the engines themselves are untouched
        same(b, state(s), "after tuning")
    path = str(tmp_path / "shapes.json")
    eatube.save_shape_cache(path, cache)
    again, t2 = eatube.tune_shape([s.engine for s in sims[::-1]],
                                  cache=eatube.load_shape_cache(path))
    assert again == best and t2 == {}


@pytest.mark.parametrize("nr", [2, 8, 16])
def test_launch_shapes_with_other_replica_counts(nr):
    """The sweep kinds with 2, 8 and 16 replicas (block maps, partials and
    energy sums by replica) equal the default shape (with cluster moves
    where there are two replicas per half)."""
    L, Lz, bc, nt, ns = 16, 32, 1, 4, 3
    betas = eatube.betas_linear(0.6, 1.6, nt)

    def sim(shape):
        x = eatube.Simulation(L, Lz, nr, nt, ns, bc=bc, betas=betas,
                              seeds=(71, 72, 73, 74), job_id=5)
        if shape:
            x.engine.set_shape(*shape)
        x.steps(2, cluster=nr >= 4)  # a Houdayer pair needs 2 replicas per half
        return state(x)

    want = sim(None)
    for shp in ((32, 512, 0, 1), (64, 1024, 1, 0), (64, 256, 2, 1), (64, 512, 3, 0),
                (64, 512, 4, 1)):
        say("nr", nr, "shape", shp)
        same(want, sim(shp), "nr %d shape %s" % (nr, shp))


def test_tiled_sweeps_at_the_widest_rows():
    """L=64: the tile (256 + 2*128 half sites from an even origin) needs two
    16-byte chunks per thread; kind 2 equals the default shape."""
    L, Lz, nr, nt, ns = 64, 16, 4, 3, 2
    betas = eatube.betas_linear(0.6, 1.6, nt)

    def sim(shape):
        x = eatube.Simulation(L, Lz, nr, nt, ns, bc=0, betas=betas,
                              seeds=(81, 82, 83, 84), job_id=6)
        if shape:
            x.engine.set_shape(*shape)
        x.steps(2, cluster=True)
        return state(x)

    want = sim(None)
    for shp in ((64, 1024, 2, 0), (64, 512, 2, 1)):
        say("L=64 shape", shp)
        same(want, sim(shp), "L=64 shape %s" % (shp,))


def test_full_48_tube_mix():
    """The batched workload at its stated size: 48 tubes, L from {8, 16,
    24}, Lz from {L, 4L, 8L}, half PBC and half OBC, 4 replicas, 24
    temperatures in [0.6, 1.6], per-tube seeds, sweep + PT + Houdayer per
    step.  After 3 steps in one batch every tube equals the same tube run
    alone; the largest PBC and OBC tubes also equal the oracle."""
    import bench
    tubes = bench.mixed_tubes(48)
    nt = 24
    specs = [(L, Lz, bc, nt, 1) for (L, Lz, bc) in tubes]
    say("48 tubes:", sorted(set((L, Lz) for L, Lz, _ in tubes)))
    plain = [make(s, 100 + k) for k, s in enumerate(specs)]
    twins = [make(s, 100 + k) for k, s in enumerate(specs)]
    batch = eatube.Batch([t.engine for t in twins])
    batch.run([t.schedule(3, cluster=True, frec_trw=1) for t in twins])
    for p in plain:
        p.steps(3, cluster=True, frec_trw=1)
    for k, (p, t) in enumerate(zip(plain, twins)):
        same(state(p), state(t), "tube %d %s" % (k, specs[k]))
    batch.close()
    picks = []
    for bc in (0, 1):
        k = max((k for k in range(48) if specs[k][2] == bc),
                key=lambda k: (specs[k][0] ** 2 * specs[k][1], -k))
        picks.append(k)
    orc.set_threads(8)
    try:
        for k in picks:
            L, Lz, bc, nt_, ns = specs[k]
            say("oracle tube", k, specs[k])
            o = orc.Sim(L, Lz, 4, nt_, ns, bc=bc,
                        betas=orc.betas_linear(0.6, 1.6, nt_),
                        seeds=SEEDS(100 + k), job_id=100 + k)
            o.steps(3, cluster=True)
            t = twins[k]
            assert np.array_equal(t.spins(), o.spins()), "tube %d spins" % k
            assert np.array_equal(t.engine.download_perms().ravel(),
                                  o.field("wctb")), "tube %d perms" % k
    finally:
        orc.set_threads(1)
