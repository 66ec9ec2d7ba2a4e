"""The oracle's OpenMP loops (samples, PT chains, cluster instances) give the
same trajectory on any thread count: the multi-core CPU baseline computes
the same thing as the one-core one."""
import numpy as np

import orc


def run(threads):
    orc.set_threads(threads)
    try:
        o = orc.Sim(8, 32, 4, 6, 5, bc=1, betas=orc.betas_linear(0.6, 1.6, 6),
                    seeds=(5, 6, 7, 8))
        o.steps(3, cluster=True)
        return {k: np.array(o.field(k), copy=True)
                for k in ("wctb", "esum", "ener", "acc", "raw")} | {
                    "spins": np.array(o.spins(), copy=True)}
    finally:
        orc.set_threads(1)


def test_thread_count_does_not_change_results():
    a, b = run(1), run(4)
    for k in a:
        assert np.array_equal(a[k], b[k]), k
