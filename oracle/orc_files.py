"""The reference program's measurement files, restated: evol.txt,
aceptancePT.txt and history_b.dat as write_measures and write_history write
them (PROGRAM/src/io.c:307-397).

TEST INFRASTRUCTURE ONLY: the tests format the oracle's values with these
functions and compare the bytes with what the simulation program wrote.
Pinned against the files the reference's own write_measures /
write_history produced from fixed arrays in the fixture harness
(oracle/ref_harness.c, tests/golden/ref_*.npz keys evol_*, aceptancePT_*,
history_*).
"""


def evol_line(ib, avg):
    """One bin's line of R%02d/evol.txt: the bin, then the average energy
    of every beta as "%14.10g " (io.c:336-343)."""
    return ("%d " % ib) + "".join("%14.10g " % float(x) for x in avg) + "\n"


def acceptance_block(ib, betas, acc, raw, att):
    """One bin's block of R%02d/aceptancePT.txt: "#bin", then per pair of
    neighbouring betas: beta, acc/att, raw/att, att as "%g" (io.c:345-356).
    """
    out = "#%d\n" % ib
    for b in range(len(betas) - 1):
        out += "%g %g %g %g\n" % (float(betas[b]), float(acc[b]) / float(att[b]),
                                  float(raw[b]) / float(att[b]), float(att[b]))
    return out


def history_append(old, ib, rows):
    """R%02d/history_b.dat after write_history(ib) (io.c:365-397): the file
    must hold at least ib * maxit_TRW * NBETAS bytes; the bin's rows
    (maxit_TRW rows of NBETAS bytes) are appended to whatever it holds,
    because the stream is opened for appending and the fseek to the
    predicted size does not move where fwrite goes -- bytes of a bin that a
    resumed run repeats stay in the file."""
    rows = bytes(bytearray(int(v) for row in rows for v in row))
    predicted = ib * len(rows)  # maxit_TRW * NBETAS bytes per bin
    if len(old) < predicted:
        raise ValueError("There are less data than expected in history_b.dat")
    return bytes(old) + rows
