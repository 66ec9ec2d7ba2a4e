"""Oracle functions that exist only in the reference's GPU code, pinned by
known answers, brute force and closed forms (not by re-typing them).

- Philox4x32-10: Random123 known-answer vectors (Salmon et al., SC'11).
- LUT bisection (main.cu:1063-1078): equals a plain search over the 256
  cumulative thresholds of the LUT.
- Metropolis algebra (main.cu:1474-1506): truth table over all 2^9 inputs
  against "unsatisfied bonds + id >= 3".
- Half-sweep (main.cu:1385-1510): against a site-by-site Metropolis on the
  unpacked lattice whose bonds are those of the reference's scalar energy
  (med.c:3-44), so neighbour words, rotations and the OBC substitution are
  checked without the MSC tables.
- Labels: against a union-find over the lattice graph.
- Random words (main.cu:1095-1150): the bits of each byte must be Bernoulli
  with p = exp(-4 beta) (create_LUT.c:249-263) and the three words
  independent.  The post-Philox mix and rotation are otherwise unpinned.
- Replica pairs of the Houdayer move (main.cu:409-421): the cluster stream
  itself is pinned by the reference's fixtures; here each call must take
  exactly NR/2 - 1 draws per group, the pair must be the first two entries
  of a forward Fisher-Yates shuffle of the group driven by the draws' high
  words (a closed form: entry 0 is position j0, entry 1 is position j1
  unless j1 = j0, where it is position 0), and over many calls every ordered
  pair of distinct replicas of a group must be equally likely.
"""
import ctypes as C

import numpy as np
import pytest

import orc


def test_philox_known_answers():
    kat = [
        ((0, 0, 0, 0), (0, 0),
         (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
        ((0xffffffff,) * 4, (0xffffffff, 0xffffffff),
         (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
        ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344),
         (0xa4093822, 0x299f31d0),
         (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)),
    ]
    for ctr, key, want in kat:
        assert tuple(int(v) for v in orc.philox(ctr, key)) == want


def lut_cum(lut_beta):
    w = np.frombuffer(lut_beta[:2048].tobytes(), np.uint32).reshape(256, 2)
    return (w[:, 0].astype(np.uint64) << np.uint64(32)) | w[:, 1].astype(
        np.uint64)


@pytest.mark.parametrize("beta", [0.05, 0.3, 0.625, 1.0, 1.43, 2.0, 3.5])
def test_bisection_is_threshold_search(beta):
    lut = orc.make_lut(beta)
    cum = lut_cum(lut)
    assert np.all(np.diff(cum[:255].astype(object)) >= 0)
    rng = np.random.default_rng(int(beta * 1000))
    R = rng.integers(0, 2**63, 4000, dtype=np.uint64) * np.uint64(2) + \
        rng.integers(0, 2, 4000, dtype=np.uint64)
    # every threshold, one below it, and the extremes
    R = np.concatenate([R, cum, cum - np.uint64(1),
                        np.array([0, 2**64 - 1], np.uint64)])
    want = np.minimum(np.searchsorted(cum, R, side="right"), 255)
    got = np.array([orc.bisect(lut, int(r >> np.uint64(32)),
                               int(r & np.uint64(0xffffffff))) for r in R])
    assert np.array_equal(got, want)


def test_flip_rule_truth_table():
    f = orc.lib().orc_flip_mask
    ones = 2**64 - 1
    for m in range(512):
        bits = [(m >> k) & 1 for k in range(9)]
        unsat = sum(bits[:6])
        b0, b1, b2 = bits[6:]
        idv = b0 + b0 * b1 + b0 * b1 * b2
        args = [ones if b else 0 for b in bits]
        got = f(*args)
        assert got == (ones if unsat + idv >= 3 else 0), (bits, got)


def site_geometry(L, Lz):
    V = L * L * Lz
    s = np.arange(V)
    x, y, z = s % L, (s // L) % L, s // (L * L)
    ML, MLz = L // 4, Lz // 4
    bit = x // ML + 4 * (y // ML) + 16 * (z // MLz)
    h = ((x % ML) + ML * (y % ML) + (z % MLz) * ML * ML) // 2
    par = (x ^ y ^ z) & 1
    return x, y, z, bit, h, par


def bond_list(L, Lz, bc, J6):
    """bonds (a, b, J) of the reference Hamiltonian, med.c:3-44"""
    V = L * L * Lz
    x, y, z, _, _, par = site_geometry(L, Lz)
    s = np.arange(V)
    xp = (x + 1) % L + y * L + z * L * L
    yp = x + ((y + 1) % L) * L + z * L * L
    zp = x + y * L + ((z + 1) % Lz) * L * L
    Jx, Jy, Jz, Jmz = J6[0:V], J6[V:2 * V], J6[2 * V:3 * V], J6[5 * V:6 * V]
    A, B, W = [s, s], [xp, yp], [Jx, Jy]
    if bc == 0:
        A.append(s); B.append(zp); W.append(Jz)
    else:
        m = z < Lz - 1
        A.append(s[m]); B.append(zp[m]); W.append(Jz[m])
        m = (z == Lz - 1) & (par == 0)
        A.append(s[m]); B.append(yp[m]); W.append(Jz[m])
        m = (z == 0) & (par == 0)
        A.append(s[m]); B.append(yp[m]); W.append(Jmz[m])
    return np.concatenate(A), np.concatenate(B), np.concatenate(W)


def unsat_count(u, bonds):
    a, b, w = bonds
    val = (u[a] ^ u[b] ^ w).astype(np.int32)
    cnt = np.zeros(u.size, np.int32)
    np.add.at(cnt, a, val)
    np.add.at(cnt, b, val)
    return cnt


CASES = [(8, 16, 0), (8, 16, 1), (16, 8, 0), (16, 8, 1), (8, 24, 1),
         (16, 16, 0), (16, 16, 1)]


@pytest.mark.parametrize("L,Lz,bc", CASES)
@pytest.mark.parametrize("variant", [0, 1, 2])
def test_half_sweep_against_site_metropolis(L, Lz, bc, variant):
    if bc == 0 and variant:
        pytest.skip("variants are OBC only")
    nr, nt = 4, 3
    V, n = L * L * Lz, L * L * Lz // 128
    rng = np.random.default_rng(L * 1000 + Lz * 10 + bc + 7 * variant)
    t = orc.MscTables(L, Lz)
    uu = rng.integers(0, 2, nr * nt * V, dtype=np.uint8)
    J6 = np.zeros(6 * V, np.uint8)
    orc.lib().orc_init_J(1234 + L + Lz, L, Lz, bc, variant, 0,
                         orc.ptr(uu, C.c_uint8), orc.ptr(J6, C.c_uint8))
    Jw = np.zeros(12 * n, np.uint64)
    orc.lib().orc_pack_J(C.byref(t.t), orc.ptr(J6, C.c_uint8),
                         orc.ptr(Jw, C.c_uint64))
    ev, od = orc.pack_spins(L, Lz, nr, nt, uu)
    spin = np.concatenate([ev, od])
    # beta of clone c for replica r: a random permutation per replica
    wc = np.zeros((nt, nr), np.uint32)
    for r in range(nr):
        wc[:, r] = rng.permutation(nt)
    rw = rng.integers(0, 2**64, 3 * nt * n * nr, dtype=np.uint64,
                      endpoint=False)
    x, y, z, bit, h, par = site_geometry(L, Lz)
    bonds = bond_list(L, Lz, bc, J6)
    u = uu.reshape(nr, nt, V).copy()
    for parity in (0, 1):
        orc.lib().orc_half_sweep(C.byref(t.t), nr, nt, bc, parity,
                                 orc.ptr(spin, C.c_uint64),
                                 orc.ptr(Jw, C.c_uint64),
                                 orc.ptr(np.ascontiguousarray(wc.ravel()),
                                         C.c_uint32),
                                 orc.ptr(rw, C.c_uint64))
        tot = nt * n * nr
        for r in range(nr):
            for c in range(nt):
                beta = wc[c, r]
                idx = (beta * n + h) * nr + r
                b = [(rw[k * tot + idx] >> bit.astype(np.uint64)) &
                     np.uint64(1) for k in range(3)]
                b0, b1, b2 = (v.astype(np.int32) for v in b)
                idv = b0 + b0 * b1 + b0 * b1 * b2
                cnt = unsat_count(u[r, c], bonds)
                flip = ((cnt + idv) >= 3) & (par == parity)
                u[r, c] ^= flip.astype(np.uint8)
        got = orc.unpack_spins(L, Lz, nr, nt, spin[:tot],
                               spin[tot:]).reshape(nr, nt, V)
        assert np.array_equal(got, u), "parity %d" % parity


def union_find_labels(L, Lz, bc, q):
    V = L * L * Lz
    parent = list(range(V))

    def find(a):
        while parent[a] != a:
            parent[a] = parent[parent[a]]
            a = parent[a]
        return a
    for s in range(V):
        x, y, z = s % L, (s // L) % L, s // (L * L)
        nbs = [(x + 1) % L + y * L + z * L * L, x + ((y + 1) % L) * L + z * L * L]
        if bc == 0 or z < Lz - 1:
            nbs.append(x + y * L + ((z + 1) % Lz) * L * L)
        for t in nbs:
            if q[s] == q[t]:
                a, b = find(s), find(t)
                if a != b:
                    parent[max(a, b)] = min(a, b)
    return np.array([find(s) for s in range(V)], np.uint32)


@pytest.mark.parametrize("L,Lz,bc,p", [(8, 16, 0, 0.5), (8, 16, 1, 0.5),
                                       (8, 8, 0, 0.15), (8, 8, 1, 0.3),
                                       (16, 8, 1, 0.5)])
def test_labels_are_minimum_index_components(L, Lz, bc, p):
    rng = np.random.default_rng(L + Lz + bc)
    q = (rng.random(L * L * Lz) < p).astype(np.uint8)
    got = orc.labels(L, Lz, bc, q)
    assert np.array_equal(got, union_find_labels(L, Lz, bc, q))
    # the label of a site is a site of its cluster with the smallest index
    assert np.all(got <= np.arange(q.size))
    assert np.all(q[got] == q)


def test_random_word_bits_are_bernoulli():
    n, nr = 64, 4
    betas = np.array([1.2, 0.8, 0.5, 0.3])
    nt = len(betas)
    lut = orc.make_luts(betas)
    rng = np.random.default_rng(5)
    tot = nt * n * nr
    b = np.zeros((3, nt, n * nr * 4), np.uint64)
    for rep in range(4):
        ctr = rng.integers(0, 2**32, 4 * nr, dtype=np.uint64).astype(np.uint32)
        keys = rng.integers(0, 2**32, nr, dtype=np.uint64).astype(np.uint32)
        w = orc.rand_words(n, nr, nt, ctr, keys, lut).reshape(3, nt, n * nr)
        b[:, :, rep * n * nr:(rep + 1) * n * nr] = w
    for i, beta in enumerate(betas):
        p = np.exp(-4 * beta)
        bits = np.unpackbits(b[:, i].view(np.uint8), axis=-1).reshape(3, -1)
        N = bits.shape[1]
        sig = np.sqrt(p * (1 - p) / N)
        for k in range(3):
            assert abs(bits[k].mean() - p) < 6 * sig, (beta, k)
        # the three words are independent draws
        assert abs((bits[0] & bits[1]).mean() - p * p) < 6 * np.sqrt(p * p / N)
        # each of the 64 bit positions is unbiased
        per_pos = bits[0].reshape(-1, 64).mean(axis=0)
        npos = bits[0].size // 64
        assert np.all(np.abs(per_pos - p) < 6 * np.sqrt(p * (1 - p) / npos))


@pytest.mark.parametrize("bc", [0, 1])
def test_overlap_is_the_sum_over_unpacked_lattices(bc):
    """orc_overlap == sum_i s_i(r1) s_i(r2) over the unpacked lattices of
    replicas r1 < r2 at the clones they hold at each beta."""
    L, Lz, nr, nt, ns = 8, 16, 4, 5, 3
    V = L * L * Lz
    o = orc.Sim(L, Lz, nr, nt, ns, bc=bc, betas=orc.betas_linear(0.6, 1.6, nt),
                seeds=(5, 6, 7, 8))
    o.steps(3, cluster=True)
    got = o.overlap()
    sp = o.spins()
    w = o.field("wctb").reshape(ns, nr, nt)
    assert got.shape == (ns, nr * (nr - 1) // 2, nt)
    for s in range(ns):
        uu = orc.unpack_spins(L, Lz, nr, nt, sp[s, 0].ravel(),
                              sp[s, 1].ravel()).reshape(nr, nt, V).astype(np.int64)
        k = 0
        for r1 in range(nr):
            for r2 in range(r1 + 1, nr):
                for b in range(nt):
                    a = 2 * uu[r1, w[s, r1, b]] - 1
                    c = 2 * uu[r2, w[s, r2, b]] - 1
                    assert got[s, k, b] == int(np.sum(a * c)), (s, r1, r2, b)
                k += 1
    assert np.all(np.abs(got) <= V) and np.any(got != V)


@pytest.mark.parametrize("ptfr,houfr,mesfr", [(1, 1, 1), (2, 1, 3), (1, 3, 2),
                                              (2, 2, 2)])
def test_it_loop_counts(ptfr, houfr, mesfr):
    """The counts main.cu:149-228 fixes per `it`, for orc_sim_run: 2*ptfr
    half-sweeps per PT, houfr PTs per Houdayer, mesfr Houdayer calls, one
    energy sum per `it`, and a TRW row every frecTRW iterations holding
    which_beta_this_clon at that iteration."""
    n_it, nr, nt, ns, V = 3, 4, 5, 2, 8 * 8 * 8
    betas = orc.betas_linear(0.6, 1.6, nt)
    o = orc.Sim(8, 8, nr, nt, ns, bc=1, betas=betas, seeds=(5, 6, 7, 8))
    hist = o.run(n_it, ptfr=ptfr, houfr=houfr, mesfr=mesfr, cluster=True,
                 it0=2, frec_trw=2, rows=3)
    calls = n_it * mesfr
    assert o.field("iter")[0] == 2 * n_it * mesfr * houfr * ptfr
    assert np.all(o.field("att") == n_it * mesfr * houfr)
    assert o.field("rng_offset")[0] == calls * 2 * V * nt * ns
    # each call draws nr/2 - 1 numbers per replica group (main.cu:409-421)
    cl = orc.hq_init(8)
    for _ in range(calls * 2 * (nr // 2 - 1)):
        orc.lib().orc_hq_next(C.byref(cl))
    assert orc.hq_to_bytes(cl) == bytes(o.field("cl"))
    # it0 = 2, frec 2: iterations 2, 3, 4 -> a row at it = 3 (row 1) only;
    # it = 3 is the middle iteration, so the row is not the final state
    assert not hist[0].any() and not hist[2].any()
    assert all(sorted(hist[1, s, r]) == list(range(nt))
               for s in range(ns) for r in range(nr))
    # the sums hold one energy per `it` and beta: for a fresh run with one
    # iteration they are the energies at the end of that iteration
    o2 = orc.Sim(8, 8, nr, nt, ns, bc=1, betas=betas, seeds=(5, 6, 7, 8))
    h2 = o2.run(1, ptfr=ptfr, houfr=houfr, mesfr=mesfr, cluster=True, rows=1)
    e = o2.field("ener").reshape(ns, nr, nt)
    w = o2.field("wctb").reshape(ns, nr, nt)
    want = np.take_along_axis(e, w.astype(np.int64), axis=2)
    assert np.array_equal(o2.field("esum").reshape(ns, nr, nt), want)
    assert np.array_equal(h2[0], o2.field("wbtc").reshape(ns, nr, nt))


def _tiny_cluster_sim(nr, seed_cl):
    betas = orc.betas_linear(0.6, 1.6, 2)
    return orc.Sim(8, 8, nr, 2, 1, seeds=(5, 6, 7, seed_cl), betas=betas)


@pytest.mark.parametrize("nr", [4, 8, 16])
def test_houdayer_pairs_are_fisher_yates_heads(nr):
    """The oracle's pair choice against the cluster stream: draws per call,
    and the first two entries of the group's shuffle in closed form."""
    h = nr // 2
    s = _tiny_cluster_sim(nr, 4242)
    cl = orc.HQ.from_buffer_copy(bytes(s.field("cl")))
    for call in range(12):
        s.houdayer()
        got = s.field("last_pairs")
        for grp in range(2):
            u = [int(orc.lib().orc_hq_next(C.byref(cl))) >> 32 for _ in range(h - 1)]
            j0 = u[0] % h
            if h == 2:
                first, second = j0, 1 - j0
            else:
                j1 = 1 + u[1] % (h - 1)
                first, second = j0, (0 if j1 == j0 else j1)
            assert (got[2 * grp], got[2 * grp + 1]) == (grp * h + first, grp * h + second), \
                (nr, call, grp)
        # the call took exactly 2 (NR/2 - 1) draws of the cluster stream
        assert bytes(s.field("cl")) == orc.hq_to_bytes(cl), (nr, call)


@pytest.mark.parametrize("nr", [4, 8])
def test_houdayer_pairs_are_uniform(nr):
    """Every ordered pair of distinct replicas of a group is equally likely
    (chi-square at the 0.1% level over 4000 calls)."""
    from scipy.stats import chi2
    h = nr // 2
    s = _tiny_cluster_sim(nr, 777)
    counts = np.zeros((2, h, h), np.int64)
    ncalls = 4000
    for _ in range(ncalls):
        s.houdayer()
        got = s.field("last_pairs")
        for grp in range(2):
            r1, r2 = int(got[2 * grp]) - grp * h, int(got[2 * grp + 1]) - grp * h
            assert 0 <= r1 < h and 0 <= r2 < h and r1 != r2
            counts[grp, r1, r2] += 1
    cells = h * (h - 1)
    expect = ncalls / cells
    for grp in range(2):
        obs = counts[grp][~np.eye(h, dtype=bool)]
        stat = float(((obs - expect) ** 2 / expect).sum())
        assert stat < chi2.ppf(0.999, cells - 1), (grp, stat)
