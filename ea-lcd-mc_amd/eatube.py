"""Python binding of the EA tube engine (libeatube.so, C ABI of
include/eatube.h) and a Simulation driver that mirrors the reference
program's initialisation and step loop (PROGRAM/src/main.cu:10-230).

The engine is HIP code for gfx950; nothing here computes physics in Python.
Loading fails loudly when the library has not been built.
"""
import ctypes as C
import json
import math
import os
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libeatube.so")
LUT_BYTES = 10240
MAX_SAMPLES = 31
PBC, OBC = 0, 1


class Config(C.Structure):
    _fields_ = [("L", C.c_int32), ("Lz", C.c_int32), ("nr", C.c_int32),
                ("nt", C.c_int32), ("ns", C.c_int32), ("bc", C.c_int32),
                ("clone_threshold", C.c_int32), ("job_id", C.c_uint32),
                ("betas", C.POINTER(C.c_double)),
                ("lut", C.POINTER(C.c_uint8))]


class Counter(C.Structure):
    _fields_ = [("x", C.c_uint32), ("y", C.c_uint32), ("z", C.c_uint32),
                ("w", C.c_uint32)]


class HQ(C.Structure):
    _fields_ = [("ira", C.c_uint64 * 256), ("PR", C.c_uint64),
                ("zseed", C.c_uint64), ("final", C.c_uint64),
                ("ip", C.c_uint8), ("ip1", C.c_uint8), ("ip2", C.c_uint8),
                ("ip3", C.c_uint8), ("padding", C.c_uint8 * 4)]


class XO(C.Structure):
    _fields_ = [("s", C.c_uint64 * 4), ("t", C.c_uint64),
                ("final", C.c_uint64)]


class Schedule(C.Structure):
    _fields_ = [("n_it", C.c_int32), ("mesfr", C.c_int32),
                ("houfr", C.c_int32), ("ptfr", C.c_int32),
                ("cluster", C.c_int32), ("it0", C.c_int32),
                ("frec_trw", C.c_int32),
                ("blocks", C.POINTER(Counter)),
                ("pairs", C.POINTER(C.c_int32)),
                ("coin_seed", C.c_uint64), ("rng_offset", C.c_uint64)]


class Shape(C.Structure):
    _fields_ = [("sweep_positions", C.c_int32), ("cluster_threads", C.c_int32),
                ("sweep_kind", C.c_int32), ("sweep_xcd", C.c_int32)]


assert C.sizeof(HQ) == 2080 and C.sizeof(XO) == 48

EXPORTS = [
    "ea_last_error", "ea_device_count", "ea_create", "ea_destroy", "ea_sync",
    "ea_set_keys", "ea_upload_sample", "ea_download_spins",
    "ea_upload_perms", "ea_download_perms", "ea_upload_pt_state",
    "ea_download_pt_state", "ea_sweep", "ea_energy", "ea_overlap", "ea_pt",
    "ea_houdayer", "ea_keep_labels", "ea_download_labels", "ea_run", "ea_read_measures",
    "ea_hq_init", "ea_hq_next", "ea_seeds_from_one", "ea_init_generators",
    "ea_fill_counters", "ea_make_lut", "ea_init_sample",
    "ea_houdayer_pairs", "ea_unpack_lattice", "ea_set_timing",
    "ea_read_timing", "ea_batch_create", "ea_batch_destroy", "ea_batch_run",
    "ea_footprint", "ea_select_device", "ea_pack_lattice", "ea_conf_bytes",
    "ea_conf_encode", "ea_conf_decode", "ea_set_shape", "ea_batch_set_shape",
    "ea_sweep_noalloc",
]
KINDS = ("sweep", "energy", "pt", "houdayer", "measure")

_lib = None


class EngineError(RuntimeError):
    pass


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise EngineError("libeatube.so is not built: run "
                          "`make -C ea-lcd-mc_amd` (or __graft_entry__.build())")
    L = C.CDLL(LIB_PATH)
    vp, i32, u32, u64 = C.c_void_p, C.c_int32, C.c_uint32, C.c_uint64
    p = C.POINTER
    sig = {
        "ea_last_error": (C.c_char_p, []),
        "ea_device_count": (C.c_int, [p(C.c_int)]),
        "ea_create": (C.c_int, [p(Config), p(vp)]),
        "ea_destroy": (C.c_int, [vp]),
        "ea_sync": (C.c_int, [vp]),
        "ea_set_keys": (C.c_int, [vp, p(u32)]),
        "ea_upload_sample": (C.c_int, [vp, C.c_int, p(u64), p(u64), p(u64)]),
        "ea_download_spins": (C.c_int, [vp, C.c_int, p(u64), p(u64)]),
        "ea_upload_perms": (C.c_int, [vp, p(C.c_uint8)]),
        "ea_download_perms": (C.c_int, [vp, p(C.c_uint8)]),
        "ea_upload_pt_state": (C.c_int, [vp, p(HQ), p(C.c_double),
                                         p(C.c_double), p(C.c_double)]),
        "ea_download_pt_state": (C.c_int, [vp, p(HQ), p(C.c_double),
                                           p(C.c_double), p(C.c_double)]),
        "ea_sweep": (C.c_int, [vp, C.c_int, p(Counter)]),
        "ea_sweep_noalloc": (C.c_int, [vp, C.c_int, p(Counter)]),
        "ea_energy": (C.c_int, [vp, p(i32)]),
        "ea_overlap": (C.c_int, [vp, p(i32)]),
        "ea_pt": (C.c_int, [vp]),
        "ea_houdayer": (C.c_int, [vp, p(i32), u64, p(u64)]),
        "ea_keep_labels": (C.c_int, [vp, C.c_int]),
        "ea_download_labels": (C.c_int, [vp, p(u32)]),
        "ea_run": (C.c_int, [vp, p(Schedule), p(u64)]),
        "ea_read_measures": (C.c_int, [vp, p(C.c_double), p(C.c_uint8),
                                       C.c_int]),
        "ea_set_timing": (C.c_int, [vp, C.c_int]),
        "ea_read_timing": (C.c_int, [vp, p(C.c_double), p(C.c_int64)]),
        "ea_batch_create": (C.c_int, [p(vp), C.c_int, p(vp)]),
        "ea_batch_destroy": (C.c_int, [vp]),
        "ea_batch_run": (C.c_int, [vp, p(Schedule), p(u64)]),
        "ea_footprint": (C.c_int, [p(Config), p(u64)]),
        "ea_set_shape": (C.c_int, [vp, p(Shape)]),
        "ea_batch_set_shape": (C.c_int, [vp, p(Shape)]),
        "ea_select_device": (C.c_int, [C.c_int]),
        "ea_pack_lattice": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int,
                                      p(C.c_uint8), C.c_int, C.c_int, p(u64),
                                      p(u64)]),
        "ea_hq_init": (None, [p(HQ), u64]),
        "ea_hq_next": (u64, [p(HQ)]),
        "ea_seeds_from_one": (None, [u64, p(u64), C.c_int]),
        "ea_init_generators": (None, [u64, p(HQ), p(XO), p(u32)]),
        "ea_fill_counters": (None, [p(HQ), p(XO), C.c_int, C.c_int, C.c_int,
                                    u32, p(Counter)]),
        "ea_make_lut": (None, [C.c_double, p(C.c_uint8)]),
        "ea_init_sample": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int,
                                     C.c_int, C.c_int, C.c_int, u64, u64,
                                     p(u64), p(u64), p(u64), p(C.c_uint8)]),
        "ea_houdayer_pairs": (C.c_int, [p(HQ), C.c_int, p(i32)]),
        "ea_unpack_lattice": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int,
                                        p(u64), p(u64), C.c_int, C.c_int,
                                        p(C.c_uint8)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _ptr(a, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


# Counter blocks as numpy records (the layout of Counter), so that a call's
# blocks need no new ctypes array type per count.
COUNTER_DTYPE = np.dtype([("x", np.uint32), ("y", np.uint32), ("z", np.uint32),
                          ("w", np.uint32)])


def _counters_arg(blocks):
    """(pointer, count) of counter blocks given as a COUNTER_DTYPE array, a
    ctypes Counter array or a sequence of Counters (the pointer keeps the
    array it points into alive)."""
    if isinstance(blocks, np.ndarray):
        if blocks.dtype != COUNTER_DTYPE or not blocks.flags.c_contiguous:
            raise ValueError("counter blocks must be a contiguous COUNTER_DTYPE array")
        return _ptr(blocks, Counter), len(blocks)
    if not isinstance(blocks, C.Array):
        blocks = (Counter * len(blocks))(*blocks)
    return C.cast(blocks, C.POINTER(Counter)), len(blocks)


_pairs_f = None


def _pairs_fn():
    """ea_houdayer_pairs taking plain addresses (no ctypes pointer objects
    per call: the pairs of every step are drawn inside timed loops)."""
    global _pairs_f
    if _pairs_f is None:
        f = lib()["ea_houdayer_pairs"]  # a function object of its own
        f.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
        f.restype = C.c_int
        _pairs_f = f
    return _pairs_f


def check(rc):
    if rc != 0:
        raise EngineError("eatube error %d: %s" % (
            rc, lib().ea_last_error().decode(errors="replace")))
    return rc


# ----------------------------------------------------------- host helpers

def seeds_from_one(seed, n):
    out = np.zeros(n, np.uint64)
    lib().ea_seeds_from_one(seed, _ptr(out, C.c_uint64), n)
    return out


def hq_init(seed):
    h = HQ()
    lib().ea_hq_init(C.byref(h), seed)
    return h


def make_lut(beta):
    out = np.zeros(LUT_BYTES, np.uint8)
    lib().ea_make_lut(float(beta), _ptr(out, C.c_uint8))
    return out


def make_luts(betas):
    return np.concatenate([make_lut(b) for b in betas])


def betas_linear(tmin, tmax, nt):
    """betas = 1/T for T linearly spaced in [tmin, tmax], descending."""
    t = np.linspace(tmin, tmax, nt)
    return np.array(sorted((1.0 / t).tolist(), reverse=True))


def clone_threshold(nt, fraction=0.5):
    """lround(nt * fraction) (HoudayerXEA.cu:136)."""
    x = nt * fraction
    return int(math.floor(x + 0.5))


def init_sample(L, Lz, nr, nt, bc, variant, mattis, seed_J, seed_u):
    n = L * L * Lz // 128
    ev = np.zeros(n * nr * nt, np.uint64)
    od = np.zeros(n * nr * nt, np.uint64)
    J = np.zeros(12 * n, np.uint64)
    cp = np.zeros(6 * L * L * Lz, np.uint8)
    cnt = lib().ea_init_sample(L, Lz, nr, nt, bc, variant, mattis, seed_J,
                               seed_u, _ptr(ev, C.c_uint64),
                               _ptr(od, C.c_uint64), _ptr(J, C.c_uint64),
                               _ptr(cp, C.c_uint8))
    if cnt < 0:
        check(cnt)
    return ev, od, J, cp, cnt


def footprint(L, Lz, nr, nt, ns, bc, betas, clone_thr=None):
    """Device bytes an engine of this configuration allocates."""
    betas = np.ascontiguousarray(betas, np.float64)
    lut = np.zeros(1, np.uint8)  # not read
    ctm = clone_threshold(nt) if clone_thr is None else clone_thr
    c = Config(L, Lz, nr, nt, ns, bc, ctm, 0, _ptr(betas, C.c_double),
               _ptr(lut, C.c_uint8))
    out = C.c_uint64()
    check(lib().ea_footprint(C.byref(c), C.byref(out)))
    return out.value


def unpack_lattice(L, Lz, nr, nt, even, odd, r, c):
    out = np.zeros(L * L * Lz, np.uint8)
    check(lib().ea_unpack_lattice(L, Lz, nr, nt, _ptr(even, C.c_uint64),
                                  _ptr(odd, C.c_uint64), r, c,
                                  _ptr(out, C.c_uint8)))
    return out


def device_count():
    n = C.c_int(0)
    rc = lib().ea_device_count(C.byref(n))
    return n.value if rc == 0 else 0


# ----------------------------------------------------------------- engine

class Engine:
    """One engine: device state of ns samples of one tube."""

    def __init__(self, L, Lz, nr, nt, ns, bc, betas, lut=None,
                 clone_thr=None, job_id=0):
        self.L, self.Lz, self.nr, self.nt, self.ns, self.bc = \
            L, Lz, nr, nt, ns, bc
        self.V = L * L * Lz
        self.n = self.V // 128
        self.half = self.n * nr * nt
        self.betas = np.ascontiguousarray(betas, np.float64)
        self.lut = np.ascontiguousarray(
            make_luts(self.betas) if lut is None else lut, np.uint8)
        self.ctm = clone_threshold(nt) if clone_thr is None else clone_thr
        c = Config(L, Lz, nr, nt, ns, bc, self.ctm, job_id,
                   _ptr(self.betas, C.c_double), _ptr(self.lut, C.c_uint8))
        self.h = C.c_void_p()
        check(lib().ea_create(C.byref(c), C.byref(self.h)))

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            lib().ea_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sync(self):
        check(lib().ea_sync(self.h))

    def set_keys(self, keys):
        k = np.ascontiguousarray(keys, np.uint32)
        check(lib().ea_set_keys(self.h, _ptr(k, C.c_uint32)))

    def upload_sample(self, s, even, odd, J):
        check(lib().ea_upload_sample(
            self.h, s, _ptr(np.ascontiguousarray(even, np.uint64), C.c_uint64),
            _ptr(np.ascontiguousarray(odd, np.uint64), C.c_uint64),
            _ptr(np.ascontiguousarray(J, np.uint64), C.c_uint64)))

    def download_spins(self, s):
        ev = np.zeros(self.half, np.uint64)
        od = np.zeros(self.half, np.uint64)
        check(lib().ea_download_spins(self.h, s, _ptr(ev, C.c_uint64),
                                      _ptr(od, C.c_uint64)))
        return ev, od

    def upload_perms(self, wctb):
        w = np.ascontiguousarray(wctb, np.uint8)
        check(lib().ea_upload_perms(self.h, _ptr(w, C.c_uint8)))

    def download_perms(self):
        w = np.zeros((self.ns, self.nr, self.nt), np.uint8)
        check(lib().ea_download_perms(self.h, _ptr(w, C.c_uint8)))
        return w

    def upload_pt_state(self, states, acc=None, att=None, raw=None):
        def d(a):
            return None if a is None else _ptr(
                np.ascontiguousarray(a, np.float64), C.c_double)
        check(lib().ea_upload_pt_state(self.h, states, d(acc), d(att),
                                       d(raw)))

    def download_pt_state(self):
        st = (HQ * (self.ns * self.nr))()
        shp = (self.ns, self.nr, self.nt - 1)
        acc, att, raw = (np.zeros(shp) for _ in range(3))
        check(lib().ea_download_pt_state(
            self.h, st, _ptr(acc, C.c_double), _ptr(att, C.c_double),
            _ptr(raw, C.c_double)))
        return st, acc, att, raw

    def sweep(self, blocks):
        b, n = _counters_arg(blocks)
        check(lib().ea_sweep(self.h, n // (2 * self.nr), b))

    def sweep_noalloc(self, blocks):
        """ea_sweep through the per-GPU counter buffer (no allocation)."""
        b, n = _counters_arg(blocks)
        check(lib().ea_sweep_noalloc(self.h, n // (2 * self.nr), b))

    def energy(self):
        out = np.zeros((self.ns, self.nr, self.nt), np.int32)
        check(lib().ea_energy(self.h, _ptr(out, C.c_int32)))
        return out

    def overlap(self):
        """site overlaps [ns][replica pair r1 < r2][beta] (ea_overlap)"""
        npairs = self.nr * (self.nr - 1) // 2
        out = np.zeros((self.ns, npairs, self.nt), np.int32)
        check(lib().ea_overlap(self.h, _ptr(out, C.c_int32)))
        return out

    def pt(self):
        check(lib().ea_pt(self.h))

    def houdayer(self, pairs, seed, offset):
        p = np.ascontiguousarray(pairs, np.int32)
        off = C.c_uint64(offset)
        check(lib().ea_houdayer(self.h, _ptr(p, C.c_int32), seed,
                                C.byref(off)))
        return off.value

    def keep_labels(self, on=True):
        check(lib().ea_keep_labels(self.h, int(on)))

    def download_labels(self):
        out = np.zeros((2, self.ns, self.ctm, self.V), np.uint32)
        check(lib().ea_download_labels(self.h, _ptr(out, C.c_uint32)))
        return out

    def run(self, sched):
        off = C.c_uint64(0)
        check(lib().ea_run(self.h, C.byref(sched), C.byref(off)))
        return off.value

    def set_shape(self, sweep_positions=64, cluster_threads=1024, sweep_kind=0,
                  sweep_xcd=0):
        """Launch shape of this engine's calls (results do not change)."""
        sh = Shape(sweep_positions, cluster_threads, sweep_kind, sweep_xcd)
        check(lib().ea_set_shape(self.h, C.byref(sh)))

    def set_timing(self, on=True):
        check(lib().ea_set_timing(self.h, int(on)))

    def read_timing(self):
        """{kind: (total ms, launches)} since the last read"""
        ms = np.zeros(len(KINDS))
        n = np.zeros(len(KINDS), np.int64)
        check(lib().ea_read_timing(self.h, _ptr(ms, C.c_double),
                                   _ptr(n, C.c_int64)))
        return {k: (float(ms[i]), int(n[i])) for i, k in enumerate(KINDS)}

    def read_measures(self, rows=0):
        es = np.zeros((self.ns, self.nr, self.nt))
        hist = np.zeros((max(rows, 1), self.ns, self.nr, self.nt), np.uint8)
        check(lib().ea_read_measures(self.h, _ptr(es, C.c_double),
                                     _ptr(hist, C.c_uint8), rows))
        return es, hist[:rows]


class Batch:
    """Several engines' tubes advancing in the same launches: one step of
    every tube captured in a HIP graph and replayed (ea_batch_run)."""

    def __init__(self, engines):
        self.engines = list(engines)
        arr = (C.c_void_p * len(self.engines))(*[e.h.value for e in self.engines])
        self.h = C.c_void_p()
        check(lib().ea_batch_create(arr, len(self.engines), C.byref(self.h)))

    def run(self, schedules):
        """schedules[i] for engine i; returns the coin offsets after the run."""
        arr = (Schedule * len(schedules))(*schedules)
        off = np.zeros(len(schedules), np.uint64)
        check(lib().ea_batch_run(self.h, arr, _ptr(off, C.c_uint64)))
        return off

    def set_shape(self, sweep_positions=64, cluster_threads=1024, sweep_kind=0,
                  sweep_xcd=0):
        """Launch shape of the batch's steps, from the next run on."""
        sh = Shape(sweep_positions, cluster_threads, sweep_kind, sweep_xcd)
        check(lib().ea_batch_set_shape(self.h, C.byref(sh)))

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            lib().ea_batch_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------------------------- simulation

class Prepared:
    """The host half of a Simulation: seeds, generators, PT states and the
    packed samples (ini.c:8-101), made without touching the GPU, so it can
    run on a worker thread (the library calls release the interpreter).

    generate_seeds_from_one emits its seeds in sequence, so only as many as
    the sample list reaches are made (the reference makes 32 and picks)."""

    def __init__(self, L, Lz, nr, nt, ns, bc=PBC, seeds=(1, 2, 3, 4),
                 variant=0, mattis=0, list_samples=None):
        seed_J, seed_u, seed_MC, seed_cl = seeds
        ls = list(range(ns)) if list_samples is None else list(list_samples)
        self.list_samples = ls
        upto = max(ls) + 1
        self.seeds_J = seeds_from_one(seed_J, upto)[ls]
        self.seeds_u = seeds_from_one(seed_u, upto)[ls]
        self.seeds_MC = seeds_from_one(seed_MC, nr)
        self.cl = hq_init(seed_cl)
        self.prc = (HQ * nr)()
        self.xo = (XO * nr)()
        self.keys = np.zeros(nr, np.uint32)
        for r in range(nr):
            k = C.c_uint32()
            lib().ea_init_generators(int(self.seeds_MC[r]), C.byref(self.prc[r]),
                                     C.byref(self.xo[r]), C.byref(k))
            self.keys[r] = k.value
        self.pts = (HQ * (ns * nr))()
        for r in range(nr):
            t = seeds_from_one(int(self.seeds_MC[r]), upto)
            for i, smp in enumerate(ls):
                s32 = int(t[smp]) & 0xFFFFFFFF  # `int seed_PT`, ini.c:63,90
                if s32 >= 2**31:
                    s32 -= 2**32
                lib().ea_hq_init(C.byref(self.pts[i * nr + r]), s32 & (2**64 - 1))
        self.samples = [init_sample(L, Lz, nr, nt, bc, variant, mattis,
                                    int(self.seeds_J[i]), int(self.seeds_u[i]))
                        for i in range(ns)]


class Simulation:
    """Seeds -> generators -> samples -> engine, as main() and Init() do
    (main.cu:70-119, ini.c:8-101), then the step loop on the device."""

    def __init__(self, L, Lz, nr, nt, ns, bc=PBC, betas=None, lut=None,
                 seeds=(1, 2, 3, 4), job_id=0, variant=0, mattis=0,
                 list_samples=None, clone_thr=None, prepared=None):
        if betas is None:
            raise ValueError("betas required")
        if prepared is None:
            prepared = Prepared(L, Lz, nr, nt, ns, bc, seeds, variant, mattis,
                                list_samples)
        p = prepared
        self.L, self.Lz, self.nr, self.nt, self.ns = L, Lz, nr, nt, ns
        self.bc, self.job = bc, job_id
        self.V = L * L * Lz
        self.seed_J, self.seed_u, self.seed_MC, self.seed_cl = seeds
        self.list_samples = p.list_samples
        self.seeds_J, self.seeds_u, self.seeds_MC = p.seeds_J, p.seeds_u, p.seeds_MC
        self.cl, self.prc, self.xo, self.keys = p.cl, p.prc, p.xo, p.keys
        self.engine = Engine(L, Lz, nr, nt, ns, bc, betas, lut=lut,
                             clone_thr=clone_thr, job_id=job_id)
        self.engine.set_keys(self.keys)
        self.countJ = []
        self.couplings = []
        for i, (ev, od, J, cp, cnt) in enumerate(p.samples):
            self.engine.upload_sample(i, ev, od, J)
            self.countJ.append(cnt)
            self.couplings.append(cp)
        self.engine.upload_pt_state(p.pts)
        self.iter = 0
        self.rng_offset = 0

    def counters(self, nhalf):
        out = np.zeros(nhalf * self.nr, COUNTER_DTYPE)
        lib().ea_fill_counters(self.prc, self.xo, self.nr, self.iter, nhalf,
                               self.job, _ptr(out, Counter))
        self.iter += nhalf
        return out

    def pairs(self, ncalls):
        out = np.zeros((ncalls, 4), np.int32)
        f = _pairs_fn()
        cl, base = C.addressof(self.cl), out.ctypes.data
        for k in range(ncalls):  # plain addresses: one cheap call per draw
            rc = f(cl, self.nr, base + 16 * k)
            if rc:
                check(rc)
        return out

    def schedule(self, n_it, ptfr=1, houfr=1, mesfr=1, cluster=True, it0=0,
                 frec_trw=1):
        """Host inputs of n_it iterations: counter blocks and replica pairs
        (drawn from the host streams exactly as the reference does)."""
        blocks = self.counters(2 * n_it * mesfr * houfr * ptfr)
        pairs = self.pairs(n_it * mesfr) if cluster else np.zeros((1, 4),
                                                                   np.int32)
        sch = Schedule(n_it, mesfr, houfr, ptfr, int(bool(cluster)), it0,
                       frec_trw, _ptr(blocks, Counter), _ptr(pairs, C.c_int32),
                       self.seed_cl, self.rng_offset)
        sch._keep = (blocks, pairs)
        if cluster:
            self.rng_offset += n_it * mesfr * 2 * self.V * self.nt * self.ns
        return sch

    def steps(self, n_it, **kw):
        sch = self.schedule(n_it, **kw)
        self.engine.run(sch)
        return sch

    def spins(self):
        """[ns][2][nt][n][nr] words"""
        e = self.engine
        out = np.zeros((self.ns, 2, e.half), np.uint64)
        for i in range(self.ns):
            out[i, 0], out[i, 1] = e.download_spins(i)
        return out.reshape(self.ns, 2, self.nt, e.n, self.nr)


# -------------------------------------------------------------- autotuner

# Launch shapes the engine takes, as (sweep_positions, cluster_threads,
# sweep_kind, sweep_xcd): the sweep options, and the Houdayer workgroup
# sizes.
SWEEP_OPTIONS = ((64, 0, 0), (32, 0, 0), (64, 0, 1), (32, 0, 1), (64, 1, 0),
                 (64, 1, 1), (64, 2, 0), (64, 2, 1), (64, 3, 0), (64, 3, 1),
                 (64, 4, 0), (64, 4, 1))
CLUSTER_THREADS = (1024, 512, 256)
DEFAULT_SHAPE = (64, 1024, 0, 0)
SHAPES = tuple((p, c, k, x) for (p, k, x) in SWEEP_OPTIONS
               for c in CLUSTER_THREADS)


def tube_key(engine):
    e = engine
    return [e.L, e.Lz, e.nr, e.nt, e.ns, e.bc, e.ctm]


def shape_signature(engines, ptfr=1, houfr=1, mesfr=1, cluster=True):
    """Cache key of a tube mix: the sorted tube configurations and the step
    frequencies (the launch sequence depends on nothing else)."""
    tubes = sorted(tube_key(e) for e in engines)
    return json.dumps({"tubes": tubes, "freq": [ptfr, houfr, mesfr],
                       "cluster": int(bool(cluster))}, separators=(",", ":"))


def load_shape_cache(path):
    if path and os.path.exists(path):
        with open(path) as f:
            return json.load(f)
    return {}


def save_shape_cache(path, cache):
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(cache, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def tune_shape(engines, ptfr=1, houfr=1, mesfr=1, cluster=True, steps=8,
               cache=None):
    """Fastest launch shape for advancing `engines` together, by timing.

    The candidates run on scratch twins (engines of the same configurations,
    betas and LUTs, with samples of their own), so the engines themselves
    are not touched.  The sweep options are timed first with the default
    Houdayer workgroup, then (with cluster moves) the workgroup sizes with
    the best sweep option: each candidate gets one untimed step (plan and
    graph are rebuilt) and then `steps` timed steps of one batch of the
    twins.  Returns (shape, {shape: seconds}) with shapes as in SHAPES; with
    a `cache` dict holding the mix's signature, returns its entry without
    timing anything, and otherwise records the winner there."""
    key = shape_signature(engines, ptfr, houfr, mesfr, cluster)
    if cache is not None and key in cache:
        return tuple(cache[key]), {}
    twins = [Simulation(e.L, e.Lz, e.nr, e.nt, e.ns, bc=e.bc, betas=e.betas,
                        lut=e.lut, clone_thr=e.ctm,
                        seeds=(101 + i, 202 + i, 303 + i, 404 + i))
             for i, e in enumerate(engines)]
    batch = Batch([t.engine for t in twins])
    kw = dict(ptfr=ptfr, houfr=houfr, mesfr=mesfr, cluster=cluster)
    times = {}

    def timed(shp):
        try:
            batch.set_shape(*shp)
        except EngineError:
            return  # a tube of the mix cannot take this shape
        batch.run([t.schedule(1, **kw) for t in twins])
        sch = [t.schedule(steps, **kw) for t in twins]
        t0 = time.perf_counter()
        batch.run(sch)
        times[shp] = time.perf_counter() - t0

    try:
        for (p, k, x) in SWEEP_OPTIONS:
            timed((p, DEFAULT_SHAPE[1], k, x))
        p, _, k, x = min(times, key=times.get)
        if cluster:
            for c in CLUSTER_THREADS:
                if (p, c, k, x) not in times:
                    timed((p, c, k, x))
    finally:
        batch.close()
        for t in twins:
            t.engine.close()
    best = min(times, key=times.get)
    if cache is not None:
        cache[key] = list(best)
    return best, times
