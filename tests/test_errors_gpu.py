"""Errors of calls on a live engine, through the C ABI: each call returns its
status code with the reason in ea_last_error, and leaves the engine as it
was -- afterwards it runs exactly as an untouched twin does."""
import ctypes as C

import numpy as np
import pytest

import eatube

pytestmark = pytest.mark.gpu

EA_E_ARG, EA_E_STATE = -1, -4


def say(*a):
    print(*a, flush=True)


def sim(nr, k):
    betas = eatube.betas_linear(0.6, 1.6, 6)
    return eatube.Simulation(16, 16, nr, 6, 2, bc=1, betas=betas,
                             seeds=(70 + k, 71 + k, 72 + k, 73 + k), job_id=k)


def expect(rc, code, text):
    msg = eatube.lib().ea_last_error()
    assert rc == code, (rc, msg)
    assert text in msg, msg


def state(s):
    e = s.engine
    st, acc, att, raw = e.download_pt_state()
    return (s.spins(), e.download_perms(), bytes(st), acc, att, raw,
            e.energy())


def test_failed_calls_leave_the_engine_unchanged():
    L = eatube.lib()
    say("case: nr=4 engine, then failing calls, then steps against a twin")
    s, twin = sim(4, 0), sim(4, 0)
    h = s.engine.h
    off = C.c_uint64(0)
    for pairs, text in (((0, 0, 2, 3), b"two distinct replicas"),
                        ((0, 2, 2, 3), b"outside group"),
                        ((0, 1, 1, 3), b"outside group")):
        p = np.array(pairs, np.int32)
        rc = L.ea_houdayer(h, p.ctypes.data_as(C.POINTER(C.c_int32)), 5,
                           C.byref(off))
        expect(rc, EA_E_ARG, text)
    b = (eatube.Counter * (2 * 4))()  # never read: both calls fail first
    expect(L.ea_sweep(h, -1, b), EA_E_ARG, b"nsweeps < 0")
    expect(L.ea_sweep(h, 1, None), EA_E_ARG, b"NULL")
    for shape, text in (((48, 1024, 0, 0), b"sweep_positions"),
                        ((64, 128, 0, 0), b"cluster_threads"),
                        ((64, 1024, 5, 0), b"sweep_kind must be"),
                        ((32, 1024, 1, 0), b"takes sweep_positions 64"),
                        ((64, 1024, 0, 2), b"sweep_xcd")):
        sh = eatube.Shape(*shape)
        expect(L.ea_set_shape(h, C.byref(sh)), EA_E_ARG, text)
    s.steps(2, cluster=True)
    twin.steps(2, cluster=True)
    for a, b_ in zip(state(s), state(twin)):
        if isinstance(a, bytes):
            assert a == b_
        else:
            assert np.array_equal(a, b_)


def test_cluster_move_needs_four_replicas():
    say("case: nr=2 engine, Houdayer call")
    s = sim(2, 1)
    p = np.array((0, 1, 0, 1), np.int32)
    off = C.c_uint64(0)
    rc = eatube.lib().ea_houdayer(s.engine.h,
                                  p.ctypes.data_as(C.POINTER(C.c_int32)), 5,
                                  C.byref(off))
    expect(rc, EA_E_STATE, b"needs nr >= 4")
    s.steps(1, cluster=False)  # the engine still runs
