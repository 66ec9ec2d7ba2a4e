"""Parity of the HIP engine with the CPU oracle (bit-exact: spins, integer
energies, PT permutations, PT generator states and statistics, cluster
labels), through the C ABI, on seeded inputs; and size-independent
properties at the full workload sizes."""
import ctypes as C
import sys

import numpy as np
import pytest

import eatube
import orc

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def pair(L, Lz, nr, nt, ns, bc=0, variant=0, mattis=0, seeds=(3, 5, 7, 9),
         job=0, tmin=0.6, tmax=1.6, ctm=None):
    betas = orc.betas_linear(tmin, tmax, nt)
    say("case L=%d Lz=%d nr=%d nt=%d ns=%d bc=%d variant=%d mattis=%d" %
        (L, Lz, nr, nt, ns, bc, variant, mattis))
    o = orc.Sim(L, Lz, nr, nt, ns, bc=bc, variant=variant, mattis=mattis,
                seeds=seeds, job_id=job, betas=betas, clonetbm=ctm)
    e = eatube.Simulation(L, Lz, nr, nt, ns, bc=bc, betas=betas, seeds=seeds,
                          job_id=job, variant=variant, mattis=mattis,
                          clone_thr=ctm)
    return o, e


def assert_same_state(o, e, what=""):
    sp = e.spins()
    osp = o.spins()
    for i in range(o.ns):
        if not np.array_equal(sp[i], osp[i]):
            bad = np.argwhere(sp[i] != osp[i])[0]
            raise AssertionError("%s spins differ: sample %d at %s" %
                                 (what, i, bad))
    assert np.array_equal(e.engine.download_perms().ravel(),
                          o.field("wctb")), what + " perms"
    st, acc, att, raw = e.engine.download_pt_state()
    assert bytes(st) == bytes(o.field("pt")), what + " PT generators"
    assert np.array_equal(acc.ravel(), o.field("acc")), what + " acceptance"
    assert np.array_equal(att.ravel(), o.field("att")), what + " attempts"
    assert np.array_equal(raw.ravel(), o.field("raw")), what + " raw"
    es, _ = e.engine.read_measures()
    assert np.array_equal(es.ravel(), o.field("esum")), what + " energy sums"
    o.field("esum")[:] = 0
    o.energy()
    assert np.array_equal(e.engine.energy().ravel(), o.field("ener")), \
        what + " energies"


def test_device_present():
    assert eatube.device_count() >= 1
    try:
        import torch
        p = torch.cuda.get_device_properties(0)
        say("device:", p.name, getattr(p, "gcnArchName", ""),
            "CUs", p.multi_processor_count,
            "LDS/block", getattr(p, "shared_memory_per_block", None),
            "LDS/block optin", getattr(p, "shared_memory_per_block_optin",
                                       None),
            "LDS/CU", getattr(p, "shared_memory_per_multiprocessor", None))
    except Exception as ex:  # informational only
        say("device properties unavailable:", ex)


@pytest.mark.parametrize("case", [
    dict(L=8, Lz=16, nr=2, nt=16, ns=3),
    dict(L=8, Lz=16, nr=2, nt=5, ns=2, bc=1),
    dict(L=16, Lz=8, nr=4, nt=3, ns=2, bc=1, variant=2),
    dict(L=8, Lz=24, nr=8, nt=3, ns=2, bc=1, variant=1),
    dict(L=24, Lz=16, nr=4, nt=4, ns=2),
    dict(L=16, Lz=16, nr=16, nt=2, ns=1, bc=1),
])
def test_sweeps_match_oracle(case):
    o, e = pair(**case)
    nr = case["nr"]
    for k in (1, 3):
        blocks = e.counters(2 * k)
        e.engine.sweep(blocks)
        o.sweeps(k)
        sp, osp = e.spins(), o.spins()
        assert np.array_equal(sp, osp), "after %d more sweeps" % k
    # the energies left by the fused odd half-sweep equal a fresh MSC energy
    o.energy()
    assert np.array_equal(e.engine.energy().ravel(), o.field("ener"))
    assert nr == e.nr


def test_plain_small_config_with_pt():
    """L=8 Lz=16 PBC, 2 replicas, 16 temperatures in [0.7,1.6], PT every
    sweep, no cluster moves (the plain small configuration, shortened)."""
    o, e = pair(8, 16, 2, 16, 31, bc=0, tmin=0.7, tmax=1.6,
                seeds=(101, 202, 303, 404), job=3)
    for chunk in (1, 4, 20):
        e.steps(chunk, cluster=False)
        o.steps(chunk, cluster=False)
        assert_same_state(o, e, "after %d steps" % chunk)


@pytest.mark.parametrize("case", [
    dict(L=8, Lz=16, nr=4, nt=6, ns=3, bc=1),
    dict(L=16, Lz=16, nr=4, nt=8, ns=2, bc=0),
    dict(L=8, Lz=24, nr=8, nt=4, ns=2, bc=1, variant=1),
    dict(L=16, Lz=8, nr=4, nt=6, ns=2, bc=1, variant=2),
    dict(L=16, Lz=64, nr=4, nt=24, ns=2, bc=1, tmin=0.6, tmax=1.6),
])
def test_steps_with_houdayer_match_oracle(case):
    o, e = pair(**case)
    o.keep_labels()
    e.engine.keep_labels()
    for chunk in (1, 3):
        e.steps(chunk, cluster=True)
        o.steps(chunk, cluster=True)
        assert_same_state(o, e, "after %d steps" % chunk)
        lab = e.engine.download_labels().ravel()
        assert np.array_equal(lab, o.field("last_labels")), "labels"


@pytest.mark.parametrize("bc", [0, 1])
def test_tiled_houdayer_matches_oracle(bc):
    """V = 16*16*264 = 67584 sites: beyond 16-bit whole-tube labels, and
    32-bit labels exceed one workgroup's LDS: slab labelling + merge."""
    o, e = pair(16, 264, 4, 4, 2, bc=bc, seeds=(21, 22, 23, 24))
    o.keep_labels()
    e.engine.keep_labels()
    e.steps(2, cluster=True)
    o.steps(2, cluster=True)
    assert_same_state(o, e, "tiled")
    assert np.array_equal(e.engine.download_labels().ravel(),
                          o.field("last_labels"))


def test_mattis_ground_state_energy():
    o, e = pair(8, 16, 4, 4, 2, bc=0, mattis=1)
    E = e.engine.energy()
    assert np.all(E[:, 0, 0] == 3 * 8 * 8 * 16)


def test_samples_are_independent_and_share_random_words():
    """Samples differ only by their data: a sample uploaded twice evolves
    identically in both slots, at any slot."""
    L, Lz, nr, nt, ns = 16, 32, 4, 8, 31
    betas = orc.betas_linear(0.6, 1.6, nt)
    e = eatube.Simulation(L, Lz, nr, nt, ns, bc=1, betas=betas,
                          seeds=(8, 9, 10, 11))
    ev, od = e.engine.download_spins(4)
    ev = ev.copy()
    od = od.copy()
    J = eatube.init_sample(L, Lz, nr, nt, 1, 0, 0, int(e.seeds_J[4]),
                           int(e.seeds_u[4]))[2]
    e.engine.upload_sample(30, ev, od, J)
    w = e.engine.download_perms()
    w[30] = w[4]
    e.engine.upload_perms(w)
    st, acc, att, raw = e.engine.download_pt_state()
    st2 = (eatube.HQ * len(st))()
    C.memmove(st2, st, C.sizeof(st))
    C.memmove(C.byref(st2, 30 * nr * C.sizeof(eatube.HQ)),
              C.byref(st, 4 * nr * C.sizeof(eatube.HQ)), nr * C.sizeof(eatube.HQ))
    e.engine.upload_pt_state(st2)
    # cluster coins are indexed by sample slot, so only sweeps and PT here
    e.steps(5, cluster=False)
    sp = e.spins()
    assert np.array_equal(sp[4], sp[30])
    assert not np.array_equal(sp[4], sp[5])


# ------------------------------------------------ full-size properties

def full_size_sim(L=48, Lz=384, nr=4, nt=32, ns=31, bc=0):
    betas = orc.betas_linear(0.5, 1.6, nt)
    return eatube.Simulation(L, Lz, nr, nt, ns, bc=bc, betas=betas,
                             seeds=(48, 384, 4, 32))


def test_full_size_one_sweep_bit_exact():
    """L=48 Lz=384 PBC, 4 replicas, 32 temperatures: one sweep of two
    samples against the oracle (the random thread index passes 2^19 here)."""
    o, e = pair(48, 384, 4, 32, 2, bc=0, tmin=0.5, tmax=1.6,
                seeds=(48, 384, 4, 32))
    e.engine.sweep(e.counters(2))
    o.sweeps(1)
    assert np.array_equal(e.spins(), o.spins())
    o.energy()
    assert np.array_equal(e.engine.energy().ravel(), o.field("ener"))


@pytest.mark.parametrize("bc", [0, 1])
def test_full_size_houdayer_involution_and_conservation(bc):
    """Large tube: the cluster move conserves E(r1)+E(r2) at every beta,
    leaves other replicas and betas above the threshold alone, and applied
    twice with the same coins restores the configuration (the reference's
    TEST_HOUDAYER_ENERGY / TEST_HOUDAYER_CONF checks, main.cu:454-526)."""
    e = full_size_sim(bc=bc, ns=8)
    eng = e.engine
    e.steps(2, cluster=False)
    say("steps done")
    before = e.spins().copy()
    E0 = eng.energy()
    w = eng.download_perms()
    pairs = np.array([1, 0, 3, 2], np.int32)
    eng.keep_labels()
    off = eng.houdayer(pairs, 777, 12345)
    after = e.spins().copy()
    E1 = eng.energy()
    ctm = eng.ctm
    for s in range(e.ns):
        for grp in (0, 1):
            r1, r2 = pairs[2 * grp], pairs[2 * grp + 1]
            for b in range(e.nt):
                c1, c2 = w[s, r1, b], w[s, r2, b]
                assert E0[s, r1, c1] + E0[s, r2, c2] == \
                    E1[s, r1, c1] + E1[s, r2, c2]
                if b >= ctm:
                    assert E0[s, r1, c1] == E1[s, r1, c1]
    assert not np.array_equal(before, after)
    lab = eng.download_labels()
    eng.houdayer(pairs, 777, 12345)
    again = e.spins()
    assert np.array_equal(again, before)
    assert off == 12345 + 2 * e.V * e.nt * e.ns
    # labels of a few instances against the oracle's breadth-first search
    L, Lz = e.L, e.Lz
    for (grp, s, b) in [(0, 0, 0), (1, 7, ctm - 1), (0, 5, ctm // 2)]:
        r1, r2 = pairs[2 * grp], pairs[2 * grp + 1]
        c1, c2 = w[s, r1, b], w[s, r2, b]
        u1 = eatube.unpack_lattice(L, Lz, e.nr, e.nt, before[s, 0].ravel(),
                                   before[s, 1].ravel(), r1, c1)
        u2 = eatube.unpack_lattice(L, Lz, e.nr, e.nt, before[s, 0].ravel(),
                                   before[s, 1].ravel(), r2, c2)
        want = orc.labels(L, Lz, bc, u1 ^ u2)
        assert np.array_equal(lab[grp, s, b], want), (grp, s, b)


def test_full_size_houdayer_31_samples():
    """L=48 Lz=384 with all 31 samples: 992 cluster instances, each
    labelled in 24 slabs joined at their seams.  The labels of every
    instance equal the oracle's breadth-first minimum-index components, and
    the move is an involution."""
    from concurrent.futures import ThreadPoolExecutor
    e = full_size_sim(bc=1, ns=31)
    eng = e.engine
    e.steps(1, cluster=False)
    before = e.spins().copy()
    w = eng.download_perms()
    pairs = np.array([0, 1, 2, 3], np.int32)
    eng.keep_labels()
    eng.houdayer(pairs, 4242, 0)
    lab = eng.download_labels()
    say("labels downloaded:", lab.shape)
    L, Lz, ctm = e.L, e.Lz, eng.ctm

    def check(inst):
        grp, s, b = inst
        r1, r2 = pairs[2 * grp], pairs[2 * grp + 1]
        c1, c2 = w[s, r1, b], w[s, r2, b]
        u1 = eatube.unpack_lattice(L, Lz, e.nr, e.nt, before[s, 0].ravel(),
                                   before[s, 1].ravel(), r1, c1)
        u2 = eatube.unpack_lattice(L, Lz, e.nr, e.nt, before[s, 0].ravel(),
                                   before[s, 1].ravel(), r2, c2)
        return np.array_equal(lab[grp, s, b], orc.labels(L, Lz, 1, u1 ^ u2))

    insts = [(g, s, b) for g in (0, 1) for s in range(e.ns)
             for b in range(ctm)]
    assert len(insts) == 992
    with ThreadPoolExecutor(16) as ex:
        ok = list(ex.map(check, insts))
    bad = [i for i, k in zip(insts, ok) if not k]
    assert not bad, bad[:5]
    say("all %d instances' labels equal the oracle's" % len(insts))
    del lab
    eng.houdayer(pairs, 4242, 0)
    assert np.array_equal(e.spins(), before)


def test_full_size_step_matches_oracle():
    """One step of the large workload (L=48 Lz=384 PBC, 4 replicas, 32
    temperatures in [0.5, 1.6]; sweep + PT + Houdayer + measure) for two
    samples against the oracle: spins, PT permutations, generators and
    statistics, energy sums, energies and every cluster label."""
    o, e = pair(48, 384, 4, 32, 2, bc=0, tmin=0.5, tmax=1.6,
                seeds=(4848, 384, 44, 3232))
    o.keep_labels()
    e.engine.keep_labels()
    orc.set_threads(8)
    try:
        e.steps(1, cluster=True)
        o.steps(1, cluster=True)
    finally:
        orc.set_threads(1)
    assert_same_state(o, e, "large step")
    assert np.array_equal(e.engine.download_labels().ravel(),
                          o.field("last_labels")), "labels"


def test_full_size_pt_matches_oracle_pt():
    """Device PT at 32 temperatures and 124 chains against upd.c's
    restatement fed the same energies and generator states."""
    e = full_size_sim(L=16, Lz=64, nt=32)
    eng = e.engine
    e.steps(3, cluster=False)
    E = eng.energy()
    st, acc, att, raw = eng.download_pt_state()
    w = eng.download_perms()
    eng.pt()
    w2 = eng.download_perms()
    st2, acc2, att2, raw2 = eng.download_pt_state()
    ost = (orc.HQ * len(st)).from_buffer_copy(bytes(st))
    betas = np.ascontiguousarray(e.engine.betas)
    for ch in range(e.ns * e.nr):
        s, r = divmod(ch, e.nr)
        a = [np.ascontiguousarray(x[s, r]) for x in (acc, att, raw)]
        wc = np.ascontiguousarray(w[s, r])
        wb = np.zeros(e.nt, np.uint8)
        En = np.ascontiguousarray(E[s, r])
        orc.lib().orc_pt(C.byref(ost[ch]), e.nt, orc.ptr(betas, C.c_double),
                         orc.ptr(wc, C.c_uint8), orc.ptr(wb, C.c_uint8),
                         orc.ptr(En, C.c_int32), orc.ptr(a[0], C.c_double),
                         orc.ptr(a[1], C.c_double), orc.ptr(a[2], C.c_double))
        assert np.array_equal(wc, w2[s, r])
        assert np.array_equal(a[0], acc2[s, r])
        assert np.array_equal(a[1], att2[s, r])
        assert np.array_equal(a[2], raw2[s, r])
    assert bytes(ost) == bytes(st2)


@pytest.mark.parametrize("nt", [220, 255])
def test_pt_many_temperatures(nt):
    """220 temperatures: one PT call draws 219 numbers, so the generator's
    256-slot ring wraps within the call.  255, the most the byte-wide
    permutations allow: every lane stages four slots and the walk runs in
    four chunks of pairs, the last one partial."""
    o, e = pair(8, 8, 2, nt, 2, bc=0, tmin=0.4, tmax=2.0,
                seeds=(61, 62, 63, 64))
    for chunk in (1, 3):
        e.steps(chunk, cluster=False)
        o.steps(chunk, cluster=False)
        assert_same_state(o, e, "after %d steps" % chunk)


@pytest.mark.parametrize("case", [
    dict(L=8, Lz=16, nr=4, nt=16, ns=5, bc=1, tmin=0.7, tmax=1.6, seeds=(11, 12, 13, 14)),
    dict(L=8, Lz=8, nr=2, nt=220, ns=2, bc=0, tmin=0.4, tmax=2.0, seeds=(61, 62, 63, 64)),
    dict(L=8, Lz=8, nr=2, nt=255, ns=3, bc=1, tmin=0.4, tmax=2.0, seeds=(65, 66, 67, 68)),
])
def test_exact_pt_walk_matches_oracle(case, monkeypatch):
    """Every PT walk forced onto the exact rule pair by pair (the path a walk
    takes when a pair's swap threshold does not verify): the same swaps,
    generator states and statistics as the oracle's."""
    monkeypatch.setenv("EA_PT_EXACT_WALK", "1")
    o, e = pair(**case)
    cluster = case["nr"] >= 4
    for chunk in (1, 3):
        e.steps(chunk, cluster=cluster)
        o.steps(chunk, cluster=cluster)
        assert_same_state(o, e, "exact walk, after %d steps" % chunk)


@pytest.mark.parametrize("case", [
    dict(L=8, Lz=16, nr=4, nt=6, ns=3, bc=1),
    dict(L=16, Lz=64, nr=4, nt=24, ns=2, bc=1, tmin=0.6, tmax=1.6),
    dict(L=24, Lz=48, nr=4, nt=4, ns=2, bc=0),
])
def test_site_forest_labelling_matches_oracle(case, monkeypatch):
    """Whole-tube labelling forced onto its forest over sites (the path an
    instance with more runs than the run forest holds takes): the same
    labels, moves and energies as the oracle's."""
    monkeypatch.setenv("EA_HOU_SITE_FOREST", "1")
    o, e = pair(**case)
    o.keep_labels()
    e.engine.keep_labels()
    for chunk in (1, 2):
        e.steps(chunk, cluster=True)
        o.steps(chunk, cluster=True)
        assert_same_state(o, e, "site forest, after %d steps" % chunk)
        assert np.array_equal(e.engine.download_labels().ravel(),
                              o.field("last_labels")), "labels"


@pytest.mark.parametrize("L,Lz,nr,nt,ns,bc", [(8, 16, 4, 6, 3, 1), (16, 32, 2, 5, 2, 0),
                                               (16, 264, 4, 3, 1, 0)])
def test_overlaps_match_oracle(L, Lz, nr, nt, ns, bc):
    """ea_overlap (popcount reductions on the device) == the oracle's site
    overlaps of every replica pair at every beta, after steps with PT and
    cluster moves (when there are two replicas per half)."""
    o, e = pair(L, Lz, nr, nt, ns, bc=bc, seeds=(31, 32, 33, 34))
    cl = nr >= 4
    e.steps(3, cluster=cl)
    o.steps(3, cluster=cl)
    got = e.engine.overlap()
    want = o.overlap()
    assert got.shape == want.shape == (ns, nr * (nr - 1) // 2, nt)
    assert np.array_equal(got, want)
    assert np.all(np.abs(got) <= L * L * Lz)


@pytest.mark.parametrize("bc", [0, 1])
def test_slab_path_for_tubes_with_few_instances(bc):
    """L=24 Lz=96 fits one workgroup's LDS but not half a CU's; with one
    sample (4 instances, fewer than the CUs) it is labelled in slabs, and
    still equals the oracle (labels, spins, energies, PT)."""
    o, e = pair(24, 96, 4, 4, 1, bc=bc, seeds=(41, 42, 43, 44))
    o.keep_labels()
    e.engine.keep_labels()
    e.steps(2, cluster=True)
    o.steps(2, cluster=True)
    assert_same_state(o, e, "24x96")
    assert np.array_equal(e.engine.download_labels().ravel(),
                          o.field("last_labels"))


def _device_free():
    """free device memory, as the driver reports it for the whole device,
    asked through a HIP runtime already mapped into this process (the
    engine's; another runtime loaded beside it may see no device)"""
    paths = []
    for line in open("/proc/self/maps"):
        f = line.split()
        if len(f) >= 6 and "libamdhip64" in f[5] and f[5] not in paths:
            paths.append(f[5])
    for path in paths:
        hip = C.CDLL(path)
        free, total = C.c_size_t(), C.c_size_t()
        if hip.hipMemGetInfo(C.byref(free), C.byref(total)) == 0:
            return free.value
    pytest.fail("no HIP runtime in this process reports device memory")


@pytest.mark.parametrize("L,Lz,ns", [(24, 96, 8), (24, 96, 31), (16, 264, 4)])
def test_footprint_matches_allocation(L, Lz, ns):
    """ea_footprint (the serving loop's admission figure) agrees with what
    ea_create allocates, on both Houdayer paths: 24x96 with 8 samples (192
    instances, fewer than the CUs) is labelled in slabs, whose buffers
    (~64 MB) dominate; with 31 samples in one workgroup per instance."""
    nr, nt = 4, 24
    betas = eatube.betas_linear(0.6, 1.6, nt)
    want = eatube.footprint(L, Lz, nr, nt, ns, 1, betas)
    warm = eatube.Simulation(8, 16, nr, nt, 1, bc=1, betas=betas, seeds=(1, 2, 3, 4))
    warm.engine.sync()  # the runtime's context and code objects are loaded
    warm.engine.close()
    before = _device_free()
    sim = eatube.Simulation(L, Lz, nr, nt, ns, bc=1, betas=betas, seeds=(1, 2, 3, 4))
    sim.engine.sync()
    used = before - _device_free()
    say("footprint", want, "allocated", used)
    # allocations are rounded to pages, and each engine adds a few small
    # runtime buffers
    assert want * 0.95 - (16 << 20) <= used <= want * 1.05 + (16 << 20)
    sim.engine.close()


def test_shape_change_after_an_empty_run_sums_pending_energies():
    """ea_sweep leaves partial energies pending; ea_run with n_it = 0 leaves
    no call in flight; ea_set_shape must still sum those partials under the
    old layout before changing it, so that the next PT decides from the
    energies of the swept configuration (perms, PT generators and
    statistics against the oracle's sweep, energy and PT)."""
    o, e = pair(16, 16, 4, 8, 3, bc=1, seeds=(41, 42, 43, 44))
    e.engine.sweep(e.counters(4))  # two sweeps
    o.sweeps(2)
    sch = eatube.Schedule(0, 1, 1, 1, 0, 0, 1, None, None, 0, 0)
    off = C.c_uint64(0)
    eatube.check(eatube.lib().ea_run(e.engine.h, C.byref(sch), C.byref(off)))
    e.engine.set_shape(32, 512, 0, 1)
    e.engine.pt()
    o.energy()
    o.pt()
    assert np.array_equal(e.engine.download_perms().ravel(), o.field("wctb"))
    st, acc, att, raw = e.engine.download_pt_state()
    assert bytes(st) == bytes(o.field("pt"))
    assert np.array_equal(acc.ravel(), o.field("acc"))
    assert np.array_equal(raw.ravel(), o.field("raw"))
    assert np.array_equal(e.engine.energy().ravel(), o.field("ener"))


@pytest.mark.parametrize("L,Lz,nt,ns,bc,ptfr", [(16, 64, 6, 31, 1, 1), (16, 64, 4, 5, 0, 2),
                                                  (8, 16, 5, 31, 0, 1), (24, 24, 4, 3, 1, 2),
                                                  (16, 128, 3, 2, 0, 1)])
def test_one_launch_sweeps_match_oracle(L, Lz, nt, ns, bc, ptfr):
    """Sweep kind 3 (both half-sweeps of whole clones in one launch, the
    clones staged in LDS) against the oracle, through ea_run with cluster
    moves and through ea_sweep: n = 128 (the headline tube: 31 samples in 8
    groups of 4, so each launch after a call's first takes its random words
    from the launch before), 8 (31 samples in one group, which makes its
    own), 108 (not a multiple of 64: per-lane energy sums) and 256 (two
    samples per group)."""
    o, e = pair(L, Lz, 4, nt, ns, bc=bc, seeds=(61, 62, 63, 64 + bc))
    e.engine.set_shape(64, 512, 3, 1)
    kw = dict(ptfr=ptfr, cluster=True)
    e.steps(3, it0=0, frec_trw=1, **kw)
    o.run(3, it0=0, frec_trw=1, rows=3, **kw)
    assert_same_state(o, e, "kind 3 run")
    e.engine.sweep(e.counters(4))  # two sweeps
    o.sweeps(2)
    o.energy()
    assert np.array_equal(e.spins(), o.spins()), "kind 3 ea_sweep spins"
    assert np.array_equal(e.engine.energy().ravel(), o.field("ener")), "kind 3 energies"


@pytest.mark.parametrize("L,Lz,nt,ns,bc,ptfr", [(16, 64, 6, 31, 1, 1), (16, 64, 4, 5, 0, 2),
                                                  (24, 24, 4, 3, 1, 2), (48, 96, 4, 7, 0, 1),
                                                  (32, 64, 3, 4, 1, 1)])
def test_slab_sweeps_match_oracle(L, Lz, nt, ns, bc, ptfr):
    """Sweep kind 4 (both half-sweeps in one launch over z-slabs of the
    clones, into a second spin array swapped after every sweep) against the
    oracle, through ea_run with cluster moves and through ea_sweep with an
    odd and then an even number of sweeps: slabs of 8 planes (two of them,
    each the other's halo across the periodic ends), of 2 planes (three,
    open boundaries), two slabs of 12 planes at L=48 and two of 8 at L=32;
    odd sample counts end the sample loop in the first LDS buffer."""
    o, e = pair(L, Lz, 4, nt, ns, bc=bc, seeds=(81, 82, 83, 84 + bc))
    e.engine.set_shape(64, 512, 4, 1)
    kw = dict(ptfr=ptfr, cluster=True)
    e.steps(3, it0=0, frec_trw=1, **kw)
    o.run(3, it0=0, frec_trw=1, rows=3, **kw)
    assert_same_state(o, e, "kind 4 run")
    e.engine.sweep(e.counters(2))  # one sweep: the arrays swap once more
    o.sweeps(1)
    o.energy()
    assert np.array_equal(e.spins(), o.spins()), "kind 4 ea_sweep spins"
    assert np.array_equal(e.engine.energy().ravel(), o.field("ener")), "kind 4 energies"
    e.engine.set_shape(64, 512, 2, 0)  # another kind continues from the current array
    e.steps(1, **kw)
    o.run(1, rows=1, **kw)
    assert_same_state(o, e, "kind 2 after kind 4")


def test_slab_sweep_refused_where_no_slab_fits():
    """An 8x16 tube (4 MSC planes) has no slab thickness whose halo planes
    stay distinct: sweep kind 4 is refused with the reason, and the engine
    keeps its shape and state."""
    o, e = pair(8, 16, 4, 4, 2, seeds=(91, 92, 93, 94))
    before = e.spins().copy()
    with pytest.raises(eatube.EngineError, match="slab thickness"):
        e.engine.set_shape(64, 512, 4, 0)
    e.steps(1, cluster=True)
    o.steps(1, cluster=True)
    assert not np.array_equal(before, e.spins())
    assert_same_state(o, e, "after the refused shape")


def test_full_size_slab_sweep_bit_exact():
    """L=48 Lz=384 PBC, 4 replicas, 32 temperatures: one kind-4 sweep
    (eight slabs of 12 planes) of three samples against the oracle, then a
    step with PT and the cluster move."""
    o, e = pair(48, 384, 4, 32, 3, bc=0, tmin=0.5, tmax=1.6,
                seeds=(48, 384, 4, 33))
    e.engine.set_shape(64, 1024, 4, 1)
    e.engine.sweep(e.counters(2))
    orc.set_threads(8)
    try:
        o.sweeps(1)
        o.energy()
        assert np.array_equal(e.spins(), o.spins())
        assert np.array_equal(e.engine.energy().ravel(), o.field("ener"))
        e.steps(1, cluster=True)
        o.steps(1, cluster=True)
    finally:
        orc.set_threads(1)
    assert_same_state(o, e, "large step, kind 4")


@pytest.mark.parametrize("ptfr,houfr,mesfr", [(2, 1, 1), (1, 2, 1), (1, 1, 2),
                                              (2, 2, 2)])
@pytest.mark.parametrize("bc", [0, 1])
def test_schedules_match_oracle(ptfr, houfr, mesfr, bc):
    """ea_run's loop nesting (main.cu:149-228): mesfr x { houfr x { ptfr
    sweeps, PT }, Houdayer }, energy sums and TRW history, against the
    oracle's it loop: spins, perms, PT generators and statistics, energy
    sums, energies and history rows.  A second call continues the bin
    (it0 > 0) and grows the history; the rows of the first call stay."""
    o, e = pair(16, 16, 4, 8, 3, bc=bc, seeds=(51, 52, 53, 54 + bc))
    kw = dict(ptfr=ptfr, houfr=houfr, mesfr=mesfr, cluster=True)
    e.steps(5, it0=1, frec_trw=2, **kw)
    h1 = o.run(5, it0=1, frec_trw=2, rows=3, **kw)
    assert_same_state(o, e, "first call")  # reads and clears the sums
    _, eh = e.engine.read_measures(rows=3)
    assert np.array_equal(eh, h1), "history rows of the first call"
    e.steps(4, it0=6, frec_trw=2, **kw)
    h2 = o.run(4, it0=6, frec_trw=2, rows=5, **kw)
    assert_same_state(o, e, "second call")
    _, eh = e.engine.read_measures(rows=5)
    assert np.array_equal(eh[:3], h1), "rows of the first call kept"
    assert np.array_equal(eh[3:], h2[3:]), "rows of the second call"
