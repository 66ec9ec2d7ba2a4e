"""ctypes access to the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, where it is the checker (or the timed CPU
restatement), never the thing the engine runs.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
LUT_BYTES = 10240
MAXNSAMPLES = 32


class HQ(C.Structure):
    _fields_ = [("ira", C.c_uint64 * 256), ("PR", C.c_uint64),
                ("zseed", C.c_uint64), ("final", C.c_uint64),
                ("ip", C.c_uint8), ("ip1", C.c_uint8), ("ip2", C.c_uint8),
                ("ip3", C.c_uint8), ("padding", C.c_uint8 * 4)]


class XO(C.Structure):
    _fields_ = [("s", C.c_uint64 * 4), ("t", C.c_uint64),
                ("final", C.c_uint64)]


class Tables(C.Structure):
    _fields_ = [("L", C.c_int), ("Lz", C.c_int), ("n", C.c_int),
                ("V", C.c_int), ("neig", C.POINTER(C.c_int)),
                ("rot", C.POINTER(C.c_uint8)),
                ("white", C.POINTER(C.c_uint32)),
                ("black", C.POINTER(C.c_uint32))]


class Cfg(C.Structure):
    _fields_ = [("L", C.c_int), ("Lz", C.c_int), ("nr", C.c_int),
                ("nt", C.c_int), ("ns", C.c_int), ("bc", C.c_int),
                ("variant", C.c_int), ("mattis", C.c_int),
                ("clonetbm", C.c_int), ("job_id", C.c_uint32),
                ("seed_J", C.c_uint64), ("seed_u", C.c_uint64),
                ("seed_MC", C.c_uint64), ("seed_cl", C.c_uint64),
                ("list_samples", C.c_int * MAXNSAMPLES),
                ("betas", C.POINTER(C.c_double)),
                ("lut", C.POINTER(C.c_uint8))]


assert C.sizeof(HQ) == 2080 and C.sizeof(XO) == 48

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        L = C.CDLL(LIB_PATH)
        vp, u8p = C.c_void_p, C.POINTER(C.c_uint8)
        u32p, u64p = C.POINTER(C.c_uint32), C.POINTER(C.c_uint64)
        sig = {
            "orc_hq_init": (None, [C.POINTER(HQ), C.c_uint64]),
            "orc_hq_next": (C.c_uint64, [C.POINTER(HQ)]),
            "orc_xo_next": (C.c_uint64, [C.POINTER(XO)]),
            "orc_init_generators": (None, [C.POINTER(HQ), C.POINTER(XO),
                                           u32p, C.c_uint64]),
            "orc_seeds_from_one": (None, [C.c_uint64, u64p, C.c_int]),
            "orc_fill_bits": (None, [C.POINTER(XO), C.POINTER(HQ), C.c_int,
                                     C.c_int, C.c_int, C.c_uint32, u32p]),
            "orc_philox": (None, [u32p, u32p, u32p]),
            "orc_bisect": (C.c_uint32, [u8p, C.c_uint32, C.c_uint32]),
            "orc_make_lut": (None, [C.c_double, u8p]),
            "orc_rand_words": (None, [C.c_int, C.c_int, C.c_int, u32p, u32p,
                                      u8p, u64p]),
            "orc_msc_tables": (C.c_int, [C.c_int, C.c_int, C.POINTER(Tables)]),
            "orc_free_tables": (None, [C.POINTER(Tables)]),
            "orc_pack_spins": (None, [C.c_int, C.c_int, C.c_int, C.c_int, u8p,
                                      u64p, u64p]),
            "orc_unpack_spins": (None, [C.c_int, C.c_int, C.c_int, C.c_int,
                                        u64p, u64p, u8p]),
            "orc_pack_J": (None, [C.POINTER(Tables), u8p, u64p]),
            "orc_init_u": (None, [C.c_uint64, C.c_int, C.c_int, C.c_int, u8p]),
            "orc_init_J": (C.c_int, [C.c_uint64, C.c_int, C.c_int, C.c_int,
                                     C.c_int, C.c_int, u8p, u8p]),
            "orc_half_sweep": (None, [C.POINTER(Tables), C.c_int, C.c_int,
                                      C.c_int, C.c_int, u64p, u64p, u32p,
                                      u64p]),
            "orc_flip_mask": (C.c_uint64, [C.c_uint64] * 9),
            "orc_energy_msc": (None, [C.POINTER(Tables), C.c_int, C.c_int,
                                      C.c_int, u64p, u64p,
                                      C.POINTER(C.c_int32)]),
            "orc_energy_scalar": (C.c_int, [C.c_int, C.c_int, C.c_int, u8p,
                                            u8p]),
            "orc_pt": (None, [C.POINTER(HQ), C.c_int, C.POINTER(C.c_double),
                              u8p, u8p, C.POINTER(C.c_int32),
                              C.POINTER(C.c_double), C.POINTER(C.c_double),
                              C.POINTER(C.c_double)]),
            "orc_labels": (None, [C.c_int, C.c_int, C.c_int, u8p, u32p]),
            "orc_coin": (C.c_int, [C.c_uint64, C.c_uint64]),
            "orc_sim_create": (vp, [C.POINTER(Cfg)]),
            "orc_sim_destroy": (None, [vp]),
            "orc_sim_keep_labels": (None, [vp, C.c_int]),
            "orc_sim_field": (vp, [vp, C.c_char_p, C.POINTER(C.c_size_t)]),
            "orc_sim_sweeps": (None, [vp, C.c_int]),
            "orc_sim_energy": (None, [vp]),
            "orc_sim_pt": (None, [vp]),
            "orc_sim_houdayer": (None, [vp]),
            "orc_overlap": (None, [vp, C.POINTER(C.c_int32)]),
            "orc_sim_steps": (None, [vp, C.c_int, C.c_int, C.c_int, C.c_int]),
            "orc_sim_run": (None, [vp] + [C.c_int] * 7 + [u8p, C.c_int]),
            "orc_set_threads": (None, [C.c_int]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def ptr(a, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


def hq_to_bytes(h):
    return bytes(C.string_at(C.addressof(h), C.sizeof(HQ)))


def hq_init(seed):
    h = HQ()
    lib().orc_hq_init(C.byref(h), seed)
    return h


def seeds_from_one(seed, n):
    out = np.zeros(n, np.uint64)
    lib().orc_seeds_from_one(seed, ptr(out, C.c_uint64), n)
    return out


def philox(ctr, key):
    c = np.array(ctr, np.uint32)
    k = np.array(key, np.uint32)
    o = np.zeros(4, np.uint32)
    lib().orc_philox(ptr(c, C.c_uint32), ptr(k, C.c_uint32),
                     ptr(o, C.c_uint32))
    return o


def make_lut(beta):
    out = np.zeros(LUT_BYTES, np.uint8)
    lib().orc_make_lut(beta, ptr(out, C.c_uint8))
    return out


def make_luts(betas):
    return np.concatenate([make_lut(b) for b in betas])


def bisect(lut_beta, msb, lsb):
    return lib().orc_bisect(ptr(lut_beta, C.c_uint8), msb, lsb)


def rand_words(n, nr, nt, ctr, keys, lut):
    out = np.zeros(3 * n * nr * nt, np.uint64)
    ctr = np.ascontiguousarray(ctr, np.uint32)
    keys = np.ascontiguousarray(keys, np.uint32)
    lib().orc_rand_words(n, nr, nt, ptr(ctr, C.c_uint32),
                         ptr(keys, C.c_uint32), ptr(lut, C.c_uint8),
                         ptr(out, C.c_uint64))
    return out


class MscTables:
    def __init__(self, L, Lz):
        self.t = Tables()
        if lib().orc_msc_tables(L, Lz, C.byref(self.t)):
            raise ValueError("bad lattice")
        n = self.t.n
        self.n = n
        self.neig = np.ctypeslib.as_array(self.t.neig, (12 * n,)).copy()
        self.rot = np.ctypeslib.as_array(self.t.rot, (2 * n,)).copy()
        self.white = np.ctypeslib.as_array(self.t.white, (n,)).copy()
        self.black = np.ctypeslib.as_array(self.t.black, (n,)).copy()

    def __del__(self):
        try:
            lib().orc_free_tables(C.byref(self.t))
        except Exception:
            pass


def pack_spins(L, Lz, nr, nt, uu):
    n = L * L * Lz // 128
    ev = np.zeros(n * nr * nt, np.uint64)
    od = np.zeros(n * nr * nt, np.uint64)
    uu = np.ascontiguousarray(uu, np.uint8)
    lib().orc_pack_spins(L, Lz, nr, nt, ptr(uu, C.c_uint8),
                         ptr(ev, C.c_uint64), ptr(od, C.c_uint64))
    return ev, od


def unpack_spins(L, Lz, nr, nt, even, odd):
    V = L * L * Lz
    uu = np.zeros(nr * nt * V, np.uint8)
    even = np.ascontiguousarray(even, np.uint64)
    odd = np.ascontiguousarray(odd, np.uint64)
    lib().orc_unpack_spins(L, Lz, nr, nt, ptr(even, C.c_uint64),
                           ptr(odd, C.c_uint64), ptr(uu, C.c_uint8))
    return uu


def energy_scalar(L, Lz, bc, u, J6):
    u = np.ascontiguousarray(u, np.uint8)
    J6 = np.ascontiguousarray(J6, np.uint8)
    return lib().orc_energy_scalar(L, Lz, bc, ptr(u, C.c_uint8),
                                   ptr(J6, C.c_uint8))


def labels(L, Lz, bc, q):
    q = np.ascontiguousarray(q, np.uint8)
    out = np.zeros(L * L * Lz, np.uint32)
    lib().orc_labels(L, Lz, bc, ptr(q, C.c_uint8), ptr(out, C.c_uint32))
    return out


def betas_linear(tmin, tmax, nt):
    """betas = 1/T, T linearly spaced in [tmin, tmax], sorted descending."""
    t = np.linspace(tmin, tmax, nt)
    return np.array(sorted((1.0 / t).tolist(), reverse=True))


def set_threads(n):
    """Threads of the oracle's sample/chain/instance loops (results do not
    depend on it); 1 unless set."""
    lib().orc_set_threads(int(n))


class Sim:
    """Full reference-algorithm simulation on the CPU."""

    FIELDS = {"spin": np.uint64, "J": np.uint64, "J6": np.uint8,
              "wctb": np.uint8, "wbtc": np.uint8, "ener": np.int32,
              "acc": np.float64, "att": np.float64, "raw": np.float64,
              "esum": np.float64,
              "pt": np.uint8, "prc": np.uint8, "xo": np.uint8,
              "keys": np.uint32, "seeds_J": np.uint64, "seeds_u": np.uint64,
              "seeds_MC": np.uint64, "cl": np.uint8, "iter": np.int32,
              "rng_offset": np.uint64, "lut": np.uint8, "countJ": np.int32,
              "last_pairs": np.int32, "last_labels": np.uint32}

    def __init__(self, L, Lz, nr, nt, ns, bc=0, variant=0, mattis=0,
                 seeds=(1, 2, 3, 4), job_id=0, betas=None, lut=None,
                 clonetbm=None, list_samples=None):
        self.L, self.Lz, self.nr, self.nt, self.ns = L, Lz, nr, nt, ns
        self.bc = bc
        self.V = L * L * Lz
        self.n = self.V // 128
        self.betas = np.ascontiguousarray(betas, np.float64)
        if clonetbm is None:
            clonetbm = int(np.floor(nt * 0.5 + 0.5))
        self.clonetbm = clonetbm
        c = Cfg()
        c.L, c.Lz, c.nr, c.nt, c.ns = L, Lz, nr, nt, ns
        c.bc, c.variant, c.mattis, c.clonetbm = bc, variant, mattis, clonetbm
        c.job_id = job_id
        c.seed_J, c.seed_u, c.seed_MC, c.seed_cl = seeds
        ls = list(range(ns)) if list_samples is None else list(list_samples)
        for i, v in enumerate(ls):
            c.list_samples[i] = v
        c.betas = ptr(self.betas, C.c_double)
        self._lut = None
        if lut is not None:
            self._lut = np.ascontiguousarray(lut, np.uint8)
            c.lut = ptr(self._lut, C.c_uint8)
        self.cfg = c
        self.h = lib().orc_sim_create(C.byref(c))
        if not self.h:
            raise ValueError("oracle sim init failed")

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_sim_destroy(self.h)
            self.h = None

    def field(self, name):
        nb = C.c_size_t()
        p = lib().orc_sim_field(self.h, name.encode(), C.byref(nb))
        dt = np.dtype(self.FIELDS[name])
        if not p or nb.value == 0:
            return np.zeros(0, dt)
        buf = (C.c_uint8 * nb.value).from_address(p)
        return np.frombuffer(buf, dtype=dt)  # a view into the sim

    def keep_labels(self, on=True):
        lib().orc_sim_keep_labels(self.h, int(on))

    def sweeps(self, k):
        lib().orc_sim_sweeps(self.h, k)

    def energy(self):
        lib().orc_sim_energy(self.h)

    def pt(self):
        lib().orc_sim_pt(self.h)

    def houdayer(self):
        lib().orc_sim_houdayer(self.h)

    def overlap(self):
        """site overlaps [ns][pair r1 < r2][beta] (orc_overlap)"""
        npairs = self.nr * (self.nr - 1) // 2
        out = np.zeros((self.ns, npairs, self.nt), np.int32)
        lib().orc_overlap(self.h, ptr(out, C.c_int32))
        return out

    def steps(self, nsteps, ptfr=1, houfr=1, cluster=True):
        lib().orc_sim_steps(self.h, nsteps, ptfr, houfr, int(cluster))

    def run(self, n_it, ptfr=1, houfr=1, mesfr=1, cluster=True, it0=0,
            frec_trw=1, rows=0):
        """n_it iterations of the `it` loop (orc_sim_run); returns the TRW
        history rows [rows][ns][nr][nt] it wrote (zero where none)."""
        hist = np.zeros((max(rows, 1), self.ns, self.nr, self.nt), np.uint8)
        lib().orc_sim_run(self.h, n_it, mesfr, houfr, ptfr, int(cluster), it0,
                          frec_trw, ptr(hist, C.c_uint8), rows)
        return hist[:rows]

    def spins(self):
        """spin words as [ns][2][nt][n][nr]"""
        return self.field("spin").reshape(self.ns, 2, self.nt, self.n,
                                          self.nr)
