"""ea_sweep_noalloc, the sweep entry point that allocates nothing (its
counter blocks go through a static per-GPU device array), against ea_sweep
on a twin engine: bit-exact spins, energies and PT state over two calls in a
row, the second smaller; a call beyond the buffer is refused and changes
nothing; engines of another size share the buffer one after the other."""
import ctypes as C
import sys

import numpy as np
import pytest

import eatube

pytestmark = pytest.mark.gpu

EA_E_ARG = -1


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def twins(L, Lz, nr, nt, ns, bc, seeds):
    betas = eatube.betas_linear(0.6, 1.6, nt)
    return [eatube.Simulation(L, Lz, nr, nt, ns, bc=bc, betas=betas,
                              seeds=seeds, job_id=5) for _ in range(2)]


def same(a, b, what):
    assert np.array_equal(a.spins(), b.spins()), what + ": spins"
    assert np.array_equal(a.engine.energy(), b.engine.energy()), \
        what + ": energies"
    assert np.array_equal(a.engine.download_perms(),
                          b.engine.download_perms()), what + ": perms"
    sa, sb = a.engine.download_pt_state(), b.engine.download_pt_state()
    assert bytes(sa[0]) == bytes(sb[0]), what + ": PT generators"


@pytest.mark.parametrize("L,Lz,nr,nt,ns,bc", [(16, 64, 4, 24, 3, 1),
                                              (8, 16, 2, 16, 31, 0)])
def test_noalloc_sweep_equals_sweep(L, Lz, nr, nt, ns, bc):
    say("case noalloc L=%d Lz=%d nr=%d nt=%d ns=%d bc=%d" %
        (L, Lz, nr, nt, ns, bc))
    a, b = twins(L, Lz, nr, nt, ns, bc, (61, 62, 63, 64))
    for nsw in (4096 // (2 * nr), 3):
        say("  call of %d sweeps" % nsw)
        a.engine.sweep(a.counters(2 * nsw))
        b.engine.sweep_noalloc(b.counters(2 * nsw))
        same(a, b, "after %d sweeps" % nsw)
        # PT from the partials either call left
        a.engine.pt()
        b.engine.pt()
        same(a, b, "after PT")
    # one sweep too many for the buffer: refused, nothing changes
    too = 4096 // (2 * nr) + 1
    blocks = (eatube.Counter * (too * 2 * nr))()
    rc = eatube.lib().ea_sweep_noalloc(b.engine.h, too, blocks)
    assert rc == EA_E_ARG
    assert b"shared buffer" in eatube.lib().ea_last_error()
    # both engines go on alike through the allocating entry points
    a.steps(2, cluster=nr >= 4)
    b.steps(2, cluster=nr >= 4)
    same(a, b, "after steps")


def test_engines_of_two_sizes_share_the_buffer_in_turn():
    say("case noalloc two sizes in turn")
    a, b = twins(16, 32, 4, 8, 2, 0, (71, 72, 73, 74))
    c, d = twins(8, 24, 8, 6, 3, 1, (81, 82, 83, 84))
    for nsw in (7, 2):
        a.engine.sweep(a.counters(2 * nsw))
        c.engine.sweep(c.counters(2 * nsw))
        b.engine.sweep_noalloc(b.counters(2 * nsw))
        b.engine.sync()  # calls on one GPU must not overlap
        d.engine.sweep_noalloc(d.counters(2 * nsw))
        d.engine.sync()
    same(a, b, "16x32")
    same(c, d, "8x24")
