"""The CPU oracle against fixtures written by the reference's own host C.

Fixtures (tests/golden/ref_*.npz) come from PROGRAM/src/{ini,io,upd,med}.c and
random_generator.cu compiled in-container (tests/golden/make_golden.py).
These pin the oracle's host generators, seeding chain, Init_u/Init_J (PBC,
OBC independent/weak/strong, Mattis), MSC packing and tables, _fill_bits,
scalar energy and Parallel_Tempering.
"""
import ctypes as C
import hashlib
import json
import os

import numpy as np

import orc
from conftest import GOLDEN


def make_sim(g):
    m = g.meta
    return orc.Sim(m["L"], m["Lz"], m["nr"], m["nt"], m["nbits"], bc=m["bc"],
                   variant=m["variant"], mattis=m["mattis"],
                   seeds=tuple(m["seeds"]), job_id=m["job"], betas=g.betas)


def test_seeds_and_generators(golden):
    s = make_sim(golden)
    golden.check("seeds_J", s.field("seeds_J"))
    golden.check("seeds_u", s.field("seeds_u"))
    golden.check("seeds_MC", s.field("seeds_MC"))
    golden.check("keys", s.field("keys"))
    golden.check("prc", s.field("prc"))
    golden.check("xo", s.field("xo"))
    golden.check("pt_init", s.field("pt"))
    golden.check("countJ", s.field("countJ"))


def test_cluster_stream_and_seed_lists(golden):
    s = make_sim(golden)
    cl = orc.HQ.from_buffer_copy(bytes(s.field("cl")))
    golden.check("cl_init", orc.hq_to_bytes(cl))
    draws = np.array([orc.lib().orc_hq_next(C.byref(cl)) for _ in range(64)],
                     np.uint64)
    golden.check("cl_draws", draws)
    lists = np.concatenate([orc.seeds_from_one(sd, 32)
                            for sd in (1, 12345, 0xDEADBEEFCAFE)])
    golden.check("seed_lists", lists)


def test_spins_couplings_packing(golden):
    s = make_sim(golden)
    m = golden.meta
    sp = s.spins()
    J6 = s.field("J6").reshape(m["nbits"], -1)
    Jw = s.field("J").reshape(m["nbits"], -1)
    for i in range(m["nbits"]):
        golden.check("even_%d" % i, sp[i, 0])
        golden.check("odd_%d" % i, sp[i, 1])
        golden.check("J6_%d" % i, J6[i])
        golden.check("couplings_%d" % i, J6[i])
        golden.check("Jw_%d" % i, Jw[i])


def test_msc_tables(golden):
    m = golden.meta
    t = orc.MscTables(m["L"], m["Lz"])
    golden.check("neig", t.neig.astype(np.int32))
    golden.check("rotate", t.rot)
    golden.check("white", t.white)
    golden.check("black", t.black)


def test_pack_unpack_round_trip(golden):
    m = golden.meta
    s = make_sim(golden)
    sp = s.spins()
    uu = orc.unpack_spins(m["L"], m["Lz"], m["nr"], m["nt"], sp[0, 0], sp[0, 1])
    ev, od = orc.pack_spins(m["L"], m["Lz"], m["nr"], m["nt"], uu)
    assert np.array_equal(ev, sp[0, 0].ravel())
    assert np.array_equal(od, sp[0, 1].ravel())


def test_scalar_and_msc_energy(golden):
    """med.c energies of the initial lattices; the MSC energy must agree."""
    m = golden.meta
    L, Lz, nr, nt, ns = m["L"], m["Lz"], m["nr"], m["nt"], m["nbits"]
    V = L * L * Lz
    s = make_sim(golden)
    sp = s.spins()
    J6 = s.field("J6").reshape(ns, 6 * V)
    es = np.zeros((ns, nr, nt), np.int32)
    for i in range(ns):
        uu = orc.unpack_spins(L, Lz, nr, nt, sp[i, 0], sp[i, 1]).reshape(
            nr, nt, V)
        for r in range(nr):
            for c in range(nt):
                es[i, r, c] = orc.energy_scalar(L, Lz, m["bc"], uu[r, c], J6[i])
    golden.check("escalar", es)
    s.energy()
    assert np.array_equal(s.field("ener").reshape(ns, nr, nt), es)
    if m["mattis"]:
        # replica 0 / clone 0 is the gauge ground state of the Mattis model
        assert np.all(es[:, 0, 0] == 3 * V)


def test_fill_bits(golden):
    m = golden.meta
    s = make_sim(golden)
    nr = m["nr"]
    prc = (orc.HQ * nr).from_buffer_copy(bytes(s.field("prc")))
    xo = (orc.XO * nr).from_buffer_copy(bytes(s.field("xo")))
    iters = np.frombuffer(golden.raw("fill_iters"), np.int32)
    out = np.zeros((len(iters), nr, 4), np.uint32)
    buf = np.zeros(4 * nr, np.uint32)
    for k, it in enumerate(iters):
        orc.lib().orc_fill_bits(xo, prc, nr, int(it), k & 1, m["job"],
                                orc.ptr(buf, C.c_uint32))
        out[k] = buf.reshape(nr, 4)
    golden.check("fill_blocks", out)


def replay_pt(golden, s, nsteps=25):
    """Parallel_Tempering as the fixture harness ran it: nsteps steps from
    the initial generators, with the scalar energies perturbed by the
    harness's rule.  Returns the end state and the per-step records."""
    m = golden.meta
    nr, nt, ns = m["nr"], m["nt"], m["nbits"]
    pts = (orc.HQ * (ns * nr)).from_buffer_copy(bytes(s.field("pt")))
    z = golden.z
    es = np.frombuffer(golden.raw("escalar"), np.int32) if golden.has(
        "escalar") and ("raw:escalar" in z) else None
    if es is None:
        s.energy()
        es = s.field("ener").copy()
    es = es.reshape(ns, nr, nt)
    wctb = np.tile(np.arange(nt, dtype=np.uint8), (ns, nr, 1))
    wbtc = wctb.copy()
    acc = np.zeros((ns, nr, nt - 1))
    att = np.zeros_like(acc)
    raw = np.zeros_like(acc)
    perms = np.zeros((nsteps, ns, nr, nt), np.uint8)
    energies = np.zeros((nsteps, ns, nr, nt), np.int32)
    betas = np.ascontiguousarray(golden.betas)
    for st in range(nsteps):
        for i in range(ns):
            for r in range(nr):
                c = np.arange(nt)
                e = es[i, r] + 4 * ((st * 7 + c * 13 + r * 3 + i) % 41 - 20)
                e = np.ascontiguousarray(e, np.int32)
                energies[st, i, r] = e
                w1 = np.ascontiguousarray(wctb[i, r])
                w2 = np.ascontiguousarray(wbtc[i, r])
                a1 = np.ascontiguousarray(acc[i, r])
                a2 = np.ascontiguousarray(att[i, r])
                a3 = np.ascontiguousarray(raw[i, r])
                orc.lib().orc_pt(C.byref(pts[i * nr + r]), nt,
                                 orc.ptr(betas, C.c_double),
                                 orc.ptr(w1, C.c_uint8), orc.ptr(w2, C.c_uint8),
                                 orc.ptr(e, C.c_int32),
                                 orc.ptr(a1, C.c_double),
                                 orc.ptr(a2, C.c_double),
                                 orc.ptr(a3, C.c_double))
                wctb[i, r], wbtc[i, r] = w1, w2
                acc[i, r], att[i, r], raw[i, r] = a1, a2, a3
        perms[st] = wctb
    return dict(pts=pts, wctb=wctb, wbtc=wbtc, acc=acc, att=att, raw=raw,
                perms=perms, energies=energies, escalar=es)


def test_parallel_tempering(golden):
    s = make_sim(golden)
    r = replay_pt(golden, s)
    acc, att, raw = r["acc"], r["att"], r["raw"]
    golden.check("pt_energies", r["energies"])
    golden.check("pt_perm", r["perms"])
    golden.check("pt_acc", np.concatenate([acc.ravel(), raw.ravel(),
                                           att.ravel()]))
    golden.check("pt_final", bytes(r["pts"]))
    golden.check("wbtc", r["wbtc"])
    # some swaps were accepted and some rejected
    assert 0 < raw.sum() < att.sum()


def test_lut_generator_matches_shipped_file():
    info = json.load(open(os.path.join(GOLDEN, "lut_nb40.json")))
    betas = np.array(info["betas"], np.float64)
    assert len(betas) == 40
    blob = betas.tobytes() + orc.make_luts(betas).tobytes()
    assert len(blob) == info["bytes"]
    assert blob[320:384].hex() == info["head_hex"]
    assert hashlib.sha256(blob).hexdigest() == info["sha256"]
