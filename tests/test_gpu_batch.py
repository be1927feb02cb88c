"""Batches of same-size instances (st_batch_*): each slot must give the
same bits as the same instance run alone through st_march, whether its
neighbours are running, finished or replaced, and whether the batch step
is replayed from a captured HIP graph or launched pass by pass."""
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

st_amd = pytest.importorskip("st_amd")


def _log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _ic(shape, seed, k0=4.0, mt=0.3):
    return st_amd.ic_isotropic(*shape, k0=k0, urms=1.0, mt=mt, seed=seed)


def _solo(shape, Q, Re, nsteps, **kw):
    s = st_amd.Solver(*shape, Re=Re, **kw)
    try:
        s.upload(Q)
        t, st, dts = s.march(nsteps)
        return s.download(), t, st
    finally:
        s.close()


@pytest.mark.parametrize("shape,kw", [((32, 32, 32), {}), ((48, 16, 32), {}),
                                       ((64, 16, 16), {"nonperiodic_x": 1, "period": (2.0, 2 * np.pi, 2 * np.pi)})])
def test_batch_slots_match_solo(shape, kw):
    _log("batch vs solo", shape, kw)
    res = [180.0, 260.0, 400.0]
    steps = [4, 2, 4]
    Qs = [_ic(shape, seed=11 + i, k0=3.0 + i) for i in range(3)]
    b = st_amd.Batch(*shape, 4, Re=100.0, **kw)
    try:
        for i in range(3):
            b.load(i, Qs[i], res[i], steps[i])
        b.march(5)                            # more than any target: finished slots idle
        step, t, dt, act = b.status()
        assert list(step[:3]) == steps and act[3] == 0
        for i in range(3):
            Qg, tg, sg = _solo(shape, Qs[i], res[i], steps[i], **kw)
            assert sg == steps[i]
            assert t[i] == tg
            assert np.array_equal(b.store(i), Qg), i
    finally:
        b.close()


def test_graph_replays_match_eager():
    """Each replay of the captured step, with eager work and syncs between
    replays, equals the same step launched pass by pass."""
    shape = (32, 32, 32)
    Q = _ic(shape, seed=5)
    g = st_amd.Batch(*shape, 2, Re=150.0)
    e = st_amd.Batch(*shape, 2, Re=150.0)
    try:
        for b in (g, e):
            b.load(0, Q, 150.0, 6)
            b.load(1, Q * 1.0, 300.0, 6)
        for it in range(6):
            g.march(1)
            os.environ["ST_NO_GRAPH"] = "1"
            try:
                e.march(1)
            finally:
                del os.environ["ST_NO_GRAPH"]
            g.sync()
            e.sync()
            for slot in (0, 1):
                assert np.array_equal(g.store(slot), e.store(slot)), (it, slot)
            assert np.array_equal(g.status()[1], e.status()[1])
    finally:
        g.close()
        e.close()


def test_slot_reuse_mid_run():
    """A slot released and reloaded while its neighbour keeps running."""
    shape = (32, 32, 32)
    Qa, Qb, Qc = _ic(shape, seed=1), _ic(shape, seed=2), _ic(shape, seed=3, k0=6.0)
    b = st_amd.Batch(*shape, 2, Re=200.0)
    try:
        b.load(0, Qa, 200.0, 6)
        b.load(1, Qb, 250.0, 2)
        b.march(2)
        b.sync()
        assert np.array_equal(b.store(1), _solo(shape, Qb, 250.0, 2)[0])
        b.release(1)
        b.load(1, Qc, 300.0, 4)
        b.march(4)
        step, t, dt, act = b.status()
        assert list(step) == [6, 4]
        assert np.array_equal(b.store(0), _solo(shape, Qa, 200.0, 6)[0])
        assert np.array_equal(b.store(1), _solo(shape, Qc, 300.0, 4)[0])
    finally:
        b.close()
