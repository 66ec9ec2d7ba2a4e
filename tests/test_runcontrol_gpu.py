"""Run control of the simulation program on the GPU, each condition forced
from the host (io.c:870-1004, timing.c:43-74): a stale `running` lock is
refused; changenbin.in extends or shortens the run; a tiny max_time stops
after one bin; SIGTERM during a run ends it with the error file and exit
code 1, and the flag=2 restart then equals the uninterrupted run."""
import os
import signal
import subprocess
import time

import pytest

from test_program_gpu import EA_TUBE, NBITS, NR, rep_dir, setup

pytestmark = pytest.mark.gpu


def cmd(*extra):
    return [EA_TUBE, "--nr", str(NR), "--obc", "0", str(NBITS), "betas",
            "input.in", "LUT_for_PRNG_nbits12_NB4.bin", "0"] + list(extra)


def run_rc(tmp, *extra):
    print("run", " ".join(cmd(*extra)), flush=True)
    r = subprocess.run(cmd(*extra), cwd=tmp, capture_output=True, text=True,
                       timeout=300)
    print(r.stdout[-1500:], r.stderr[-1500:], flush=True)
    return r


def idir(tmp):
    return tmp / "out" / "I000"


def confs(tmp):
    d = rep_dir(tmp, 0, 0)
    return sorted(p.name for p in d.glob("conf.*")) if d.exists() else []


def same_state(a, b):
    """conf bytes (but the start flag, bytes 28-31), evol.txt and
    history_b.dat of every (sample, replica)."""
    for bit in range(NBITS):
        for r in range(NR):
            da, db = rep_dir(a, bit, r), rep_dir(b, bit, r)
            for f in ("conf", "evol.txt", "history_b.dat"):
                x, y = (da / f).read_bytes(), (db / f).read_bytes()
                if f == "conf":
                    x, y = x[:28] + x[32:], y[:28] + y[32:]
                assert x == y, (bit, r, f)


def test_stale_lock_is_refused(tmp_path):
    setup(tmp_path, nbin=1, flag=1)
    idir(tmp_path).mkdir(parents=True)
    lock = idir(tmp_path) / "running"
    lock.write_text("otherhost 12345 1700000000 a run that died\n")
    r = run_rc(tmp_path)
    assert r.returncode == 1
    assert "running exists" in r.stderr
    assert lock.read_text().startswith("otherhost 12345")  # left alone
    assert confs(tmp_path) == []


@pytest.mark.parametrize("nbin,change", [(1, 3), (5, 2)])
def test_changenbin_sets_the_number_of_bins(tmp_path, nbin, change):
    a, b = tmp_path / "changed", tmp_path / "plain"
    a.mkdir()
    b.mkdir()
    setup(a, nbin=nbin, flag=1)
    idir(a).mkdir(parents=True)
    (idir(a) / "changenbin.in").write_text("%d\n" % change)
    r = run_rc(a)
    assert r.returncode == 0, r.stderr
    assert "nbin changed to %d" % change in r.stdout
    assert not (idir(a) / "changenbin.in").exists()
    assert confs(a) == ["conf.%06d" % k for k in range(1, change + 1)]
    setup(b, nbin=change, flag=1)
    assert run_rc(b).returncode == 0
    same_state(a, b)


def test_tiny_max_time_stops_after_one_bin(tmp_path):
    setup(tmp_path, nbin=5, flag=1)
    r = run_rc(tmp_path, "1")
    assert r.returncode == 0, r.stderr
    assert "finished at iteration 0" in r.stdout
    assert confs(tmp_path) == ["conf.000001"]
    assert not (idir(tmp_path) / "running").exists()
    assert not (idir(tmp_path) / "error").exists()


def test_sigterm_then_restart_equals_uninterrupted(tmp_path):
    a, b = tmp_path / "killed", tmp_path / "straight"
    a.mkdir()
    b.mkdir()
    setup(a, nbin=2000, flag=1, itmax=200, mesfr=1, ptfr=1)
    print("start", " ".join(cmd()), flush=True)
    p = subprocess.Popen(cmd(), cwd=a, stdout=subprocess.PIPE,
                         stderr=subprocess.PIPE, text=True)
    try:
        t0 = time.time()
        while "conf.000001" not in confs(a):  # the first bin is written
            assert p.poll() is None, p.communicate()
            assert time.time() - t0 < 120, "no bin after 120 s"
            time.sleep(0.01)
        # a signal that lands while the bin's files are written is ignored
        # (as in the reference); the bins' GPU work takes most of the time
        for _ in range(500):
            p.send_signal(signal.SIGTERM)
            try:
                p.wait(timeout=0.02)
                break
            except subprocess.TimeoutExpired:
                pass
        out, err = p.communicate(timeout=60)
    finally:
        if p.poll() is None:
            p.kill()
            p.wait()
    print(out[-800:], err[-800:], flush=True)
    assert p.returncode == 1
    assert "Aborting programma (SIGNAL=%d)" % signal.SIGTERM in err
    assert (idir(a) / "error").read_text().startswith("Received SIGTERM")
    assert not (idir(a) / "running").exists()
    done = int(confs(a)[-1][5:])
    assert done >= 1
    print("stopped after %d bins" % done, flush=True)
    # restart from the last conf for one more bin; then the same bins in one
    # uninterrupted run
    setup(a, nbin=done + 1, flag=2, itmax=200, mesfr=1, ptfr=1)
    r = run_rc(a)
    assert r.returncode == 0, r.stderr
    setup(b, nbin=done + 1, flag=1, itmax=200, mesfr=1, ptfr=1)
    assert run_rc(b).returncode == 0
    same_state(a, b)
