"""The simulation program end to end on the GPU: the reference's inputs
and output tree; its final configurations equal the library driven step by
step from the same seeds; and a run stopped after one bin and resumed from
its conf files ends byte-identical to the uninterrupted run."""
import ctypes as C
import os
import subprocess
import sys

import numpy as np
import pytest

import eatube
from conftest import ROOT
from test_conf import RunData, lib as conf_lib, ptr

pytestmark = pytest.mark.gpu

PKG = os.path.join(ROOT, "ea-lcd-mc_amd")
EA_TUBE = os.path.join(PKG, "ea_tube")
EA_LUT = os.path.join(PKG, "ea_make_lut")

L, LZ, NT, NR, NBITS = 8, 16, 4, 4, 2
SEEDS = (1357, 2468, 97531, 8642)
BETAS = [1.5, 1.2, 0.9, 0.65]


def setup(tmp, nbin, flag, itmax=4, mesfr=2, ptfr=2, houfr=1):
    (tmp / "out").mkdir(exist_ok=True)
    (tmp / "betas").write_text("".join("%.17g\n" % b for b in BETAS))
    if not (tmp / "LUT_for_PRNG_nbits12_NB4.bin").exists():
        subprocess.run([EA_LUT, "betas"], cwd=tmp, check=True, timeout=300,
                       capture_output=True)
    lines = [str(tmp / "out") + "/", nbin, 0, itmax, houfr, mesfr, ptfr, NT,
             flag, (LZ << 16) + L] + list(SEEDS)
    (tmp / "input.in").write_text("".join("%s\t\tx\n" % v for v in lines))


def run(tmp, *extra):
    cmd = [EA_TUBE, "--nr", str(NR), "--obc", "0", str(NBITS), "betas",
           "input.in", "LUT_for_PRNG_nbits12_NB4.bin", "0"] + list(extra)
    print("run", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, cwd=tmp, capture_output=True, text=True,
                       timeout=300)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0, r.stderr
    return r


def rep_dir(tmp, b, r):
    return tmp / "out" / "I000" / ("BIT%03d" % b) / ("R%02d" % r)


def decode(blob):
    n = L * L * LZ // 128 * NR * NT
    ev, od = np.zeros(n, np.uint64), np.zeros(n, np.uint64)
    w = np.zeros(NT, np.uint8)
    e = np.zeros(NT, np.int32)
    d = RunData()
    return ev, od, w, e, d


def test_program_matches_library_and_resumes(tmp_path):
    a = tmp_path / "straight"
    b = tmp_path / "resumed"
    a.mkdir()
    b.mkdir()
    setup(a, nbin=2, flag=1)
    run(a)
    setup(b, nbin=1, flag=1)
    run(b)
    setup(b, nbin=2, flag=2)  # conf files present: back up and go on
    run(b)
    V = L * L * LZ
    for bit in range(NBITS):
        for r in range(NR):
            da, db = rep_dir(a, bit, r), rep_dir(b, bit, r)
            for f in ("conf", "conf.000002", "evol.txt", "history_b.dat"):
                x, y = (da / f).read_bytes(), (db / f).read_bytes()
                if f.startswith("conf"):
                    # s_data.flag (bytes 28-31) records how each run started:
                    # 1 for a new run, 0 for a backup; the rest is the state
                    assert x[28] == 1 and y[28] == 0
                    x, y = x[:28] + x[32:], y[:28] + y[32:]
                assert x == y, (bit, r, f)
            assert os.path.samefile(da / "conf", da / "conf.000002")
            assert (da / "history_b.dat").stat().st_size == 2 * 4 * NT
            assert len((da / "evol.txt").read_text().splitlines()) == 2
            assert (da / "aceptancePT.txt").read_text().count("#") == 2
        assert (a / "out" / "I000" / ("BIT%03d" % bit) / "couplings.bin"
                ).stat().st_size == 6 * V
        assert (a / "out" / "I000" / ("BIT%03d" % bit) / "seeds.in").exists()
    assert (a / "out" / "I000" / "master_seeds.in").read_text().split() == \
        [str(s) for s in SEEDS]
    assert not (a / "out" / "I000" / "running").exists()

    # the same run through the library, step by step
    sim = eatube.Simulation(L, LZ, NR, NT, NBITS, bc=eatube.OBC,
                            betas=np.array(BETAS), seeds=SEEDS, job_id=0)
    for _ in range(2):
        sim.steps(4, mesfr=2, ptfr=2, cluster=True)
    sp = sim.spins().reshape(NBITS, 2, -1)
    perms = sim.engine.download_perms()
    ener = sim.engine.energy()
    size = conf_lib().ea_conf_bytes(L, LZ, NT)
    for bit in range(NBITS):
        for r in range(NR):
            blob = np.frombuffer((rep_dir(a, bit, r) / "conf").read_bytes(),
                                 np.uint8).copy()
            assert blob.size == size
            ev, od, w, e, d = decode(blob)
            rc = conf_lib().ea_conf_decode(
                ptr(blob, C.c_uint8), blob.size, L, LZ, NR, NT, C.byref(d),
                None, None, None, None, None, None, ptr(w, C.c_uint8),
                ptr(ev, C.c_uint64), ptr(od, C.c_uint64), r, ptr(e, C.c_int32))
            assert rc == 0
            assert d.itcut == 2 and d.nbin == 2
            assert np.array_equal(w, perms[bit, r])
            assert np.array_equal(e, ener[bit, r][w])
            for c in range(NT):
                x = eatube.unpack_lattice(L, LZ, NR, NT, ev, od, r, c)
                y = eatube.unpack_lattice(L, LZ, NR, NT,
                                          np.ascontiguousarray(sp[bit, 0]),
                                          np.ascontiguousarray(sp[bit, 1]), r, c)
                assert np.array_equal(x, y), (bit, r, c)


def test_program_refuses_a_second_new_run(tmp_path):
    setup(tmp_path, nbin=1, flag=1)
    run(tmp_path)
    cmd = [EA_TUBE, "--nr", str(NR), "--obc", "0", str(NBITS), "betas",
           "input.in", "LUT_for_PRNG_nbits12_NB4.bin", "0"]
    r = subprocess.run(cmd, cwd=tmp_path, capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 1 and "There are confs" in r.stderr
    assert (tmp_path / "out" / "I000" / "error").exists()


def frec_trw(itmax):
    """frecTRW and maxit_TRW as read_input picks them (io.c:431-434)."""
    for f in range(1, itmax + 1):
        if itmax % f == 0 and itmax // f <= 512:
            return f, itmax // f


class OracleFiles:
    """The program's measurement files as the oracle's run of the same
    seeds gives them, formatted by the reference's writers (orc_files)."""

    def __init__(self, itmax, mesfr, ptfr, houfr):
        import orc
        self.o = orc.Sim(L, LZ, NR, NT, NBITS, bc=1, seeds=SEEDS, job_id=0,
                         betas=np.array(BETAS))
        self.k = (itmax, mesfr, ptfr, houfr)
        self.frec, self.maxit = frec_trw(itmax)
        self.evol, self.acc, self.hist = {}, {}, {}

    def bin(self, ib):
        import orc_files
        itmax, mesfr, ptfr, houfr = self.k
        o = self.o
        rows = o.run(itmax, ptfr=ptfr, houfr=houfr, mesfr=mesfr, cluster=True,
                     it0=0, frec_trw=self.frec, rows=self.maxit)
        avg = o.field("esum").reshape(NBITS, NR, NT) / itmax
        o.field("esum")[:] = 0
        acc, att, raw = (o.field(k).reshape(NBITS, NR, NT - 1)
                         for k in ("acc", "att", "raw"))
        for bit in range(NBITS):
            for r in range(NR):
                key = (bit, r)
                self.evol[key] = self.evol.get(key, "") + \
                    orc_files.evol_line(ib, avg[bit, r])
                self.acc[key] = self.acc.get(key, "") + \
                    orc_files.acceptance_block(ib, BETAS, acc[bit, r],
                                               raw[bit, r], att[bit, r])
                self.hist[key] = orc_files.history_append(
                    self.hist.get(key, b""), ib, rows[:, bit, r])

    def check(self, tmp, accept=True):
        for bit in range(NBITS):
            for r in range(NR):
                d = rep_dir(tmp, bit, r)
                assert (d / "evol.txt").read_text() == self.evol[bit, r], \
                    (bit, r, "evol.txt")
                if accept:
                    assert (d / "aceptancePT.txt").read_text() == \
                        self.acc[bit, r], (bit, r, "aceptancePT.txt")
                assert (d / "history_b.dat").read_bytes() == \
                    self.hist[bit, r], (bit, r, "history_b.dat")


@pytest.mark.parametrize("itmax,mesfr,ptfr,houfr", [
    (4, 2, 2, 1), (3, 1, 1, 2), (2, 2, 1, 2), (1024, 1, 1, 1)])
def test_program_files_match_oracle(tmp_path, itmax, mesfr, ptfr, houfr):
    """evol.txt, aceptancePT.txt and history_b.dat of a two-bin run, byte
    for byte, against the oracle's energies, PT statistics and TRW rows
    written as the reference's write_measures / write_history write them
    (itmax = 1024 stores every second iteration: frecTRW = 2)."""
    setup(tmp_path, nbin=2, flag=1, itmax=itmax, mesfr=mesfr, ptfr=ptfr,
          houfr=houfr)
    run(tmp_path)
    want = OracleFiles(itmax, mesfr, ptfr, houfr)
    for ib in range(2):
        want.bin(ib)
    want.check(tmp_path)


def test_restart_counter_omits_houfr(tmp_path):
    """With houfr = 2 the reference restarts its half-sweep counter at
    2*itcut*itmax*mesfr*ptfr (main.cu:139), not at the count the first bin
    reached, so a resumed run continues on other random words than an
    uninterrupted one.  The program keeps that: its resumed run equals the
    oracle run whose counter is reset that way after the first bin (PT
    statistics, which the conf does not hold, restart from zero)."""
    itmax, mesfr, ptfr, houfr = 3, 1, 2, 2
    a, b = tmp_path / "resumed", tmp_path / "straight"
    a.mkdir()
    b.mkdir()
    setup(a, nbin=1, flag=1, itmax=itmax, mesfr=mesfr, ptfr=ptfr, houfr=houfr)
    run(a)
    setup(a, nbin=2, flag=2, itmax=itmax, mesfr=mesfr, ptfr=ptfr, houfr=houfr)
    run(a)
    want = OracleFiles(itmax, mesfr, ptfr, houfr)
    want.bin(0)
    o = want.o
    assert o.field("iter")[0] == 2 * itmax * mesfr * houfr * ptfr
    o.field("iter")[0] = 2 * 1 * itmax * mesfr * ptfr
    for k in ("acc", "att", "raw"):
        o.field(k)[:] = 0
    want.bin(1)
    want.check(a)
    # the uninterrupted run differs: the quirk matters at houfr > 1
    setup(b, nbin=2, flag=1, itmax=itmax, mesfr=mesfr, ptfr=ptfr, houfr=houfr)
    run(b)
    assert any((rep_dir(a, bit, r) / "evol.txt").read_bytes() !=
               (rep_dir(b, bit, r) / "evol.txt").read_bytes()
               for bit in range(NBITS) for r in range(NR))
