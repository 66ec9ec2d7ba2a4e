"""Checkpoint records (conf files) against the reference's own write_conf
(io.c:17-104) output held by the fixtures: the fixture harness wrote the
conf of sample 0 / replica 0 and of the last sample / replica after its
Init, 40 counter fills, 25 PT steps and a coin offset of 123456789, with the
initial scalar energies.  The same state is rebuilt here from the pinned
host generators and the oracle's PT, encoded by the library's conf encoder
(ea_conf_encode), and must be byte-identical.  Decoding round-trips."""
import ctypes as C

import numpy as np

import eatube
import orc
from test_oracle_golden import make_sim, replay_pt


class RunData(C.Structure):
    _fields_ = [(k, C.c_int32) for k in ("nbin", "itcut", "itmax", "houfr",
                                         "mesfr", "ptfr", "nbetas", "flag",
                                         "l", "pad_")] + \
               [(k, C.c_uint64) for k in ("seed_J", "seed_u", "seed_MC",
                                          "seed_Cluster")]


def lib():
    L = eatube.lib()
    p = C.POINTER
    L.ea_conf_bytes.restype = C.c_size_t
    L.ea_conf_bytes.argtypes = [C.c_int, C.c_int, C.c_int]
    L.ea_conf_encode.restype = C.c_int
    L.ea_conf_encode.argtypes = [
        p(RunData), C.c_int, C.c_int, C.c_int, C.c_int, p(C.c_double),
        p(eatube.HQ), p(eatube.HQ), p(eatube.HQ), p(eatube.XO), C.c_uint64,
        p(C.c_uint8), p(C.c_uint64), p(C.c_uint64), C.c_int, p(C.c_int32),
        p(C.c_uint8)]
    L.ea_conf_decode.restype = C.c_int
    L.ea_conf_decode.argtypes = [
        p(C.c_uint8), C.c_size_t, C.c_int, C.c_int, C.c_int, C.c_int,
        p(RunData), p(C.c_double), p(eatube.HQ), p(eatube.HQ), p(eatube.HQ),
        p(eatube.XO), p(C.c_uint64), p(C.c_uint8), p(C.c_uint64),
        p(C.c_uint64), C.c_int, p(C.c_int32)]
    return L


def ptr(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def harness_state(golden):
    m = golden.meta
    L, Lz, nt, ns, nr = m["L"], m["Lz"], m["nt"], m["nbits"], m["nr"]
    s = make_sim(golden)
    sp = s.spins().reshape(ns, 2, -1).copy()  # the oracle hands out views
    r = replay_pt(golden, s)
    seeds = m["seeds"]
    sJ = eatube.seeds_from_one(seeds[0], 32)[:ns]
    su = eatube.seeds_from_one(seeds[1], 32)[:ns]
    sMC = eatube.seeds_from_one(seeds[2], nr)
    prc = (eatube.HQ * nr)()
    xo = (eatube.XO * nr)()
    for k in range(nr):
        key = C.c_uint32()
        eatube.lib().ea_init_generators(int(sMC[k]), C.byref(prc[k]),
                                        C.byref(xo[k]), C.byref(key))
    out = (eatube.Counter * (40 * nr))()
    eatube.lib().ea_fill_counters(prc, xo, nr, 0, 40, m["job"], out)
    cl = eatube.hq_init(seeds[3])
    pts = (eatube.HQ * (ns * nr)).from_buffer_copy(bytes(r["pts"]))
    return dict(m=m, L=L, Lz=Lz, nt=nt, ns=ns, nr=nr, sp=sp, sJ=sJ, su=su,
                sMC=sMC, prc=prc, xo=xo, cl=cl, pts=pts, wctb=r["wctb"],
                E0=np.ascontiguousarray(r["escalar"], np.int32),
                betas=np.ascontiguousarray(golden.betas, np.float64))


def encode(st, i, rr):
    m = st["m"]
    d = RunData(1, 0, 1, 1, 1, 1, st["nt"], 1, (st["Lz"] << 16) + st["L"], 0,
                int(st["sJ"][i]), int(st["su"][i]), int(st["sMC"][rr]),
                m["seeds"][3])
    size = lib().ea_conf_bytes(st["L"], st["Lz"], st["nt"])
    out = np.zeros(size, np.uint8)
    ev = np.ascontiguousarray(st["sp"][i, 0])
    od = np.ascontiguousarray(st["sp"][i, 1])
    w = np.ascontiguousarray(st["wctb"][i, rr])
    e = np.ascontiguousarray(st["E0"][i, rr])
    rc = lib().ea_conf_encode(
        C.byref(d), st["L"], st["Lz"], st["nr"], st["nt"],
        ptr(st["betas"], C.c_double), C.byref(st["pts"][i * st["nr"] + rr]),
        C.byref(st["prc"][rr]), C.byref(st["cl"]), C.byref(st["xo"][rr]),
        123456789, ptr(w, C.c_uint8), ptr(ev, C.c_uint64),
        ptr(od, C.c_uint64), rr, ptr(e, C.c_int32), ptr(out, C.c_uint8))
    assert rc == 0
    return out


def test_conf_bytes_match_reference(golden):
    st = harness_state(golden)
    golden.check("conf_0_0", encode(st, 0, 0))
    golden.check("conf_last", encode(st, st["ns"] - 1, st["nr"] - 1))


def test_conf_decode_round_trip(golden):
    st = harness_state(golden)
    i, rr = st["ns"] - 1, 1
    blob = encode(st, i, rr)
    n = st["sp"].shape[2]
    ev = np.zeros(n, np.uint64)
    od = np.zeros(n, np.uint64)
    d = RunData()
    betas = np.zeros(st["nt"])
    pt, prc, cl = eatube.HQ(), eatube.HQ(), eatube.HQ()
    xo = eatube.XO()
    off = C.c_uint64()
    w = np.zeros(st["nt"], np.uint8)
    e = np.zeros(st["nt"], np.int32)
    rc = lib().ea_conf_decode(
        ptr(blob, C.c_uint8), blob.size, st["L"], st["Lz"], st["nr"],
        st["nt"], C.byref(d), ptr(betas, C.c_double), C.byref(pt),
        C.byref(prc), C.byref(cl), C.byref(xo), C.byref(off),
        ptr(w, C.c_uint8), ptr(ev, C.c_uint64), ptr(od, C.c_uint64), rr,
        ptr(e, C.c_int32))
    assert rc == 0
    assert d.itmax == 1 and d.l == (st["Lz"] << 16) + st["L"]
    assert d.seed_J == int(st["sJ"][i]) and off.value == 123456789
    assert np.array_equal(betas, st["betas"])
    assert bytes(pt) == bytes(st["pts"][i * st["nr"] + rr])
    assert bytes(xo) == bytes(st["xo"][rr]) and bytes(cl) == bytes(st["cl"])
    assert np.array_equal(w, st["wctb"][i, rr])
    assert np.array_equal(e, st["E0"][i, rr][w])
    # only replica rr's bits were written: compare through the unpacker
    nr, nt = st["nr"], st["nt"]
    for c in range(nt):
        a = eatube.unpack_lattice(st["L"], st["Lz"], nr, nt, ev, od, rr, c)
        b = eatube.unpack_lattice(st["L"], st["Lz"], nr, nt,
                                  np.ascontiguousarray(st["sp"][i, 0]),
                                  np.ascontiguousarray(st["sp"][i, 1]), rr, c)
        assert np.array_equal(a, b)
    # a damaged permutation or size is refused
    bad = blob.copy()
    rc = lib().ea_conf_decode(ptr(bad, C.c_uint8), bad.size - 1, st["L"],
                              st["Lz"], nr, nt, None, None, None, None, None,
                              None, None, None, None, None, rr, None)
    assert rc != 0
