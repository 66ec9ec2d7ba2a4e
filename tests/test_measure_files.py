"""The measurement files of the reference program -- evol.txt,
aceptancePT.txt and history_b.dat (io.c:307-397) -- restated in
oracle/orc_files.py, against the bytes the reference's own write_measures
and write_history wrote in the fixture harness: two bins of energy
averages and PT statistics, and the history of bins 0 and 1 followed by
bin 1 again, as a run resumed after losing bin 1's conf writes it."""
import numpy as np

import orc_files
from test_oracle_golden import make_sim, replay_pt


def harness_inputs(golden):
    m = golden.meta
    nr, nt, ns = m["nr"], m["nt"], m["nbits"]
    s = make_sim(golden)
    r = replay_pt(golden, s)
    es = r["escalar"].reshape(ns, nr * nt)
    avg = np.zeros((2, ns, nr * nt))
    for b in range(2):
        for i in range(ns):
            for k in range(nr * nt):
                e0 = float(es[i, k])
                avg[b, i, k] = e0 + ((k * 5 + i + b * 3) % 11) / 7.0 - \
                    (1e-3 * es[i, k] if b else 0.0)
    golden.check("meas_avg", avg)
    wbtc = r["wbtc"]
    rows = np.zeros((3, ns, nr, 3, nt), np.uint8)
    for call in range(3):
        for i in range(ns):
            for rr in range(nr):
                for it in range(3):
                    rows[call, i, rr, it] = (wbtc[i, rr].astype(int) + it +
                                             2 * call + rr) % nt
    golden.check("hist_rows", rows)
    return avg, r, rows


def test_measure_files_equal_the_reference_writers(golden):
    m = golden.meta
    nr, nt, ns = m["nr"], m["nt"], m["nbits"]
    avg, r, rows = harness_inputs(golden)
    for tag, (i, rr) in (("0_0", (0, 0)), ("last", (ns - 1, nr - 1))):
        evol = "".join(orc_files.evol_line(b, avg[b, i, rr * nt:(rr + 1) * nt])
                       for b in range(2))
        golden.check("evol_" + tag, evol.encode())
        acc = "".join(orc_files.acceptance_block(
            b, golden.betas, r["acc"][i, rr], r["raw"][i, rr], r["att"][i, rr])
            for b in range(2))
        golden.check("aceptancePT_" + tag, acc.encode())
        hist = b""
        for call, ib in ((0, 0), (1, 1), (2, 1)):
            hist = orc_files.history_append(hist, ib, rows[call, i, rr])
        golden.check("history_" + tag, hist)


def test_history_refuses_a_short_file():
    rows = np.zeros((3, 4), np.uint8)
    try:
        orc_files.history_append(b"\0" * 11, 1, rows)
    except ValueError as e:
        assert "less data" in str(e)
    else:
        raise AssertionError("a file shorter than the bins before is refused")
