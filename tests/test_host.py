"""Host side of the engine library (no GPU): the library loads and exports
every symbol of include/eatube.h, and its seeding, generators, sample
initialisation, counter blocks, LUT and replica-pair draws equal the
oracle's (itself pinned against the reference's host C) and the reference
fixtures."""
import ctypes as C
import hashlib
import json
import os
import re

import numpy as np
import pytest

import eatube
import orc
from conftest import GOLDEN, ROOT, Golden


def header_symbols():
    text = open(os.path.join(ROOT, "include", "eatube.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ea_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    syms = header_symbols()
    assert len(syms) >= 25
    L = eatube.lib()
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing
    assert sorted(eatube.EXPORTS) == syms


def test_errors_are_reported_not_fatal():
    # an invalid configuration fails with a message, without a GPU
    c = eatube.Config(12, 16, 4, 8, 1, 0, 4, 0, None, None)
    h = C.c_void_p()
    rc = eatube.lib().ea_create(C.byref(c), C.byref(h))
    assert rc == -1
    assert b"multiples of 8" in eatube.lib().ea_last_error()
    sh = eatube.Shape(64, 1024, 0, 0)
    assert eatube.lib().ea_set_shape(None, C.byref(sh)) == -1
    assert eatube.lib().ea_batch_set_shape(None, C.byref(sh)) == -1


def _config(nt=8, sorted_betas=True, with_tables=True, **over):
    """A valid configuration (L=16, Lz=16, 4 replicas, 8 betas, 2 samples,
    PBC, 4 cluster betas) with `over` replacing fields."""
    f = dict(L=16, Lz=16, nr=4, nt=nt, ns=2, bc=0, clone_threshold=4,
             job_id=0)
    f.update(over)
    betas = np.linspace(1.6, 0.6, max(nt, 1))
    if not sorted_betas:
        betas = betas[::-1].copy()
    keep = (betas, np.zeros(max(nt, 1) * 10240, dtype=np.uint8))
    b = keep[0].ctypes.data_as(C.POINTER(C.c_double)) if with_tables else None
    lut = keep[1].ctypes.data_as(C.POINTER(C.c_uint8)) if with_tables else None
    c = eatube.Config(f["L"], f["Lz"], f["nr"], f["nt"], f["ns"], f["bc"],
                      f["clone_threshold"], f["job_id"], b, lut)
    return c, keep


@pytest.mark.parametrize("over,msg", [
    (dict(L=12), b"multiples of 8"),
    (dict(Lz=20), b"multiples of 8"),
    (dict(nr=3), b"power of two"),
    (dict(nr=1), b"power of two"),
    (dict(nr=32), b"power of two"),
    (dict(nt=1), b"out of 2..255"),
    (dict(nt=256), b"out of 2..255"),
    (dict(ns=0), b"out of 1..31"),
    (dict(ns=32), b"out of 1..31"),
    (dict(bc=2), b"bc=2"),
    (dict(clone_threshold=-1), b"clone_threshold"),
    (dict(clone_threshold=9), b"clone_threshold"),
    (dict(job_id=1 << 14), b"14-bit"),
    (dict(with_tables=False), b"betas and lut are required"),
    (dict(sorted_betas=False), b"sorted from largest to smallest"),
])
def test_invalid_configurations_fail_with_their_reason(over, msg):
    """Every check ea_create makes before it looks for a device: the call
    returns EA_E_ARG (-1) with the reason, and no engine."""
    c, keep = _config(**over)
    h = C.c_void_p(1)
    rc = eatube.lib().ea_create(C.byref(c), C.byref(h))
    assert rc == -1
    assert h.value is None
    assert msg in eatube.lib().ea_last_error(), eatube.lib().ea_last_error()
    del keep


def test_inexact_index_quotients_are_refused():
    """L=112, Lz=144168 passes the size checks, but the half-site quotient
    by L*L/32 = 392 taken by multiply is wrong at h = 14128463 (36042
    instead of 36041): creation and the footprint query refuse the tube
    instead of letting the sweep update the wrong neighbours."""
    c, keep = _config(L=112, Lz=144168, clone_threshold=0)
    h = C.c_void_p(1)
    assert eatube.lib().ea_create(C.byref(c), C.byref(h)) == -1
    assert h.value is None
    msg = eatube.lib().ea_last_error()
    assert b"not exact" in msg and b"L*L/32" in msg, msg
    out = C.c_uint64()
    assert eatube.lib().ea_footprint(C.byref(c), C.byref(out)) == -1
    assert b"not exact" in eatube.lib().ea_last_error()
    del keep


def workload_lattices():
    """Every (L, Lz) of the plain, OBC, batched, large and serving
    workloads: 8x16, 16x64, L in {8, 16, 24} with Lz in {L, 4L, 8L},
    48x384, and L in {8, 16, 24, 32} with Lz a multiple of 8 in [L, 16L]."""
    out = {(8, 16), (16, 64), (48, 384)}
    out |= {(L, k * L) for L in (8, 16, 24) for k in (1, 4, 8)}
    out |= {(L, Lz) for L in (8, 16, 24, 32) for Lz in range(L, 16 * L + 1, 8)}
    return sorted(out)


def test_every_workload_lattice_passes_the_quotient_check():
    betas = np.array([1.6, 0.6])
    for L, Lz in workload_lattices():
        assert eatube.footprint(L, Lz, 4, 2, 1, 0, betas) > 0, (L, Lz)


@pytest.mark.parametrize("seed", [1, 12345, 0xDEADBEEFCAFE, 2**63 + 5])
def test_seed_lists_and_hq(seed):
    assert np.array_equal(eatube.seeds_from_one(seed, 32),
                          orc.seeds_from_one(seed, 32))
    a, b = eatube.hq_init(seed), orc.hq_init(seed)
    assert bytes(a) == bytes(b)
    for _ in range(300):
        assert eatube.lib().ea_hq_next(C.byref(a)) == \
            orc.lib().orc_hq_next(C.byref(b))


def test_generators_and_counters_match_oracle():
    nr, job = 4, 77
    prc_e, xo_e = (eatube.HQ * nr)(), (eatube.XO * nr)()
    prc_o, xo_o = (orc.HQ * nr)(), (orc.XO * nr)()
    keys_e, keys_o = np.zeros(nr, np.uint32), np.zeros(nr, np.uint32)
    for r in range(nr):
        k = C.c_uint32()
        eatube.lib().ea_init_generators(1000 + r, C.byref(prc_e[r]),
                                        C.byref(xo_e[r]), C.byref(k))
        keys_e[r] = k.value
        orc.lib().orc_init_generators(C.byref(prc_o[r]), C.byref(xo_o[r]),
                                      orc.ptr(keys_o[r:], C.c_uint32),
                                      1000 + r)
    assert np.array_equal(keys_e, keys_o)
    assert bytes(prc_e) == bytes(prc_o) and bytes(xo_e) == bytes(xo_o)
    nhalf, iter0 = 70, 32760
    out = (eatube.Counter * (nhalf * nr))()
    eatube.lib().ea_fill_counters(prc_e, xo_e, nr, iter0, nhalf, job, out)
    got = np.frombuffer(bytes(out), np.uint32).reshape(nhalf, nr, 4)
    buf = np.zeros(4 * nr, np.uint32)
    for k in range(nhalf):
        orc.lib().orc_fill_bits(xo_o, prc_o, nr, iter0 + k, k & 1, job,
                                orc.ptr(buf, C.c_uint32))
        assert np.array_equal(got[k], buf.reshape(nr, 4)), k


def test_counters_against_reference_fixture():
    g = Golden("ref_L8_Lz16_NB4_PBC.npz")
    nr = g.meta["nr"]
    s = orc.Sim(g.meta["L"], g.meta["Lz"], nr, g.meta["nt"], 1,
                seeds=tuple(g.meta["seeds"]), job_id=g.meta["job"],
                betas=g.betas)
    prc = (eatube.HQ * nr).from_buffer_copy(bytes(s.field("prc")))
    xo = (eatube.XO * nr).from_buffer_copy(bytes(s.field("xo")))
    iters = np.frombuffer(g.raw("fill_iters"), np.int32)
    # the fixture's first 32 entries are consecutive half-sweeps 0..31
    out = (eatube.Counter * (32 * nr))()
    eatube.lib().ea_fill_counters(prc, xo, nr, 0, 32, g.meta["job"], out)
    got = np.frombuffer(bytes(out), np.uint32).reshape(32, nr, 4)
    assert np.array_equal(iters[:32], np.arange(32))
    g.check("fill_blocks", _fixture_prefix(g, got))


def _fixture_prefix(g, got):
    """bytes of the fixture's fill_blocks with the first 32 rows replaced by
    `got`: equal to the fixture iff those rows match"""
    nr = g.meta["nr"]
    s = orc.Sim(g.meta["L"], g.meta["Lz"], nr, g.meta["nt"], 1,
                seeds=tuple(g.meta["seeds"]), job_id=g.meta["job"],
                betas=g.betas)
    prc = (orc.HQ * nr).from_buffer_copy(bytes(s.field("prc")))
    xo = (orc.XO * nr).from_buffer_copy(bytes(s.field("xo")))
    iters = np.frombuffer(g.raw("fill_iters"), np.int32)
    rows = []
    buf = np.zeros(4 * nr, np.uint32)
    for k, it in enumerate(iters):
        orc.lib().orc_fill_bits(xo, prc, nr, int(it), k & 1, g.meta["job"],
                                orc.ptr(buf, C.c_uint32))
        rows.append(got[k] if k < 32 else buf.reshape(nr, 4).copy())
    return np.array(rows)


@pytest.mark.parametrize("fname", ["ref_L8_Lz16_NB4_PBC.npz",
                                   "ref_L8_Lz16_NB4_OBC.npz",
                                   "ref_L8_Lz24_NB3_OBC_WEAK.npz",
                                   "ref_L16_Lz8_NB3_OBC_STRONG.npz",
                                   "ref_L8_Lz16_NB4_PBC_MATTIS.npz",
                                   "ref_L8_Lz16_NB4_OBC_MATTIS.npz",
                                   "ref_L16_Lz64_NB24_OBC.npz"])
def test_init_sample_against_reference(fname):
    g = Golden(fname)
    m = g.meta
    seeds_J = eatube.seeds_from_one(m["seeds"][0], 32)
    seeds_u = eatube.seeds_from_one(m["seeds"][1], 32)
    g.check("seeds_J", seeds_J[:m["nbits"]])
    g.check("seeds_u", seeds_u[:m["nbits"]])
    counts = []
    for i in range(m["nbits"]):
        ev, od, J, cp, cnt = eatube.init_sample(
            m["L"], m["Lz"], m["nr"], m["nt"], m["bc"], m["variant"],
            m["mattis"], int(seeds_J[i]), int(seeds_u[i]))
        g.check("even_%d" % i, ev)
        g.check("odd_%d" % i, od)
        g.check("Jw_%d" % i, J)
        g.check("couplings_%d" % i, cp)
        counts.append(cnt)
    g.check("countJ", np.array(counts, np.int32))


def test_unpack_matches_oracle():
    L, Lz, nr, nt = 16, 24, 4, 3
    ev, od, _, _, _ = eatube.init_sample(L, Lz, nr, nt, 1, 0, 0, 5, 6)
    uu = orc.unpack_spins(L, Lz, nr, nt, ev, od).reshape(nr, nt, -1)
    for r in range(nr):
        for c in range(nt):
            assert np.array_equal(
                eatube.unpack_lattice(L, Lz, nr, nt, ev, od, r, c), uu[r, c])


def test_lut_matches_shipped_file():
    info = json.load(open(os.path.join(GOLDEN, "lut_nb40.json")))
    betas = np.array(info["betas"], np.float64)
    blob = betas.tobytes() + eatube.make_luts(betas).tobytes()
    assert hashlib.sha256(blob).hexdigest() == info["sha256"]


@pytest.mark.parametrize("nr", [4, 8, 16])
def test_houdayer_pairs_match_oracle(nr):
    betas = orc.betas_linear(0.6, 1.6, 4)
    s = orc.Sim(8, 8, nr, 4, 1, seeds=(11, 12, 13, 99), betas=betas)
    cl = eatube.hq_init(99)
    for _ in range(5):
        s.houdayer()
        want = s.field("last_pairs")
        got = np.zeros(4, np.int32)
        eatube.check(eatube.lib().ea_houdayer_pairs(
            C.byref(cl), nr, got.ctypes.data_as(C.POINTER(C.c_int32))))
        assert np.array_equal(got, want)
        assert got[0] != got[1] and got[2] != got[3]
        assert max(got[:2]) < nr // 2 <= min(got[2:])


def test_pt_seeds_are_sign_extended_int():
    """Init_Random stores the PT seed in an int (ini.c:63,90): the product
    Simulation must derive the same PT generators as the reference."""
    g = Golden("ref_L8_Lz16_NB4_PBC.npz")
    m = g.meta
    nr, ns = m["nr"], m["nbits"]
    seeds_MC = eatube.seeds_from_one(m["seeds"][2], nr)
    g.check("seeds_MC", seeds_MC)
    pts = (eatube.HQ * (ns * nr))()
    for r in range(nr):
        t = eatube.seeds_from_one(int(seeds_MC[r]), 32)
        for i in range(ns):
            s32 = int(t[i]) & 0xFFFFFFFF
            if s32 >= 2**31:
                s32 -= 2**32
            eatube.lib().ea_hq_init(C.byref(pts[i * nr + r]),
                                    s32 & (2**64 - 1))
    g.check("pt_init", bytes(pts))


def test_seed_prefix_property():
    """generate_seeds_from_one emits seeds in sequence: asking for fewer
    gives the same leading seeds (the engine makes only those it uses)."""
    for seed in (1, 987654321, 2**63 + 11):
        full = eatube.seeds_from_one(seed, 32)
        for k in (1, 3, 17):
            assert np.array_equal(eatube.seeds_from_one(seed, k), full[:k])
