"""Groups: batches of 48^3, 64^3, 96^3 and 128^3 stepped with one launch per
pass over all of them (st_group_*).

Bar: each slot bit-identical to the same slot in a plain batch of its size
whose passes all take the layouts a group uses (M/2 points per lane,
256-thread tiles), step counters and times equal; and the 48^3 slot within
1e-10 norm-wise of the oracle after 3 steps (the multi-step bar of the
single-instance parity tests).
"""
import sys

import numpy as np
import pytest

import oracle as O
from fields import relerr

pytestmark = pytest.mark.gpu

st_amd = pytest.importorskip("st_amd")


def _log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


KW = dict(Pr=0.7, hyper_param=1)
# slots per size, and the Reynolds numbers of the slots
PLAN = {48: [60.0, 250.0], 64: [120.0], 96: [90.0, 300.0], 128: [200.0]}


def _ic(n, seed):
    return st_amd.ic_isotropic(n, n, n, k0=4.0, urms=1.0, mt=0.3, seed=seed)


def _batches(narrow):
    out = {}
    for n, res in PLAN.items():
        b = st_amd.Batch(n, n, n, len(res), **KW)
        if narrow:
            b.set_pass_configs(st_amd.narrow_configs(b))
        for s, re in enumerate(res):
            b.load(s, _ic(n, 100 * n + s), re, nsteps=50)
        out[n] = b
    return out


def _close(bs):
    for b in bs.values():
        b.close()


def _same(grp, ref, what):
    for n, res in PLAN.items():
        sg, tg, _, ag = grp[n].status()
        sr, tr, _, ar = ref[n].status()
        assert (sg == sr).all() and (ag == ar).all(), (what, n, sg, sr)
        assert (tg == tr).all(), (what, n, tg, tr)
        for s in range(len(res)):
            if not ag[s] and not ar[s]:
                continue
            a, b = grp[n].store(s), ref[n].store(s)
            assert np.array_equal(a, b), (what, n, s, np.abs(a - b).max())


def test_group_matches_batches():
    _log("group: sizes", sorted(PLAN))
    grp, ref = _batches(False), _batches(True)
    g = st_amd.Group(grp.values())
    try:
        g.march(3)
        g.sync()
        for b in ref.values():
            b.march(3)
            b.sync()
        _log("  3 steps")
        _same(grp, ref, "3 steps")
        # a slot leaves one member and another member empties: the steps
        # after run other launch grids (another captured graph)
        grp[48].release(1)
        ref[48].release(1)
        grp[128].release(0)
        ref[128].release(0)
        g.march(2)
        for b in ref.values():
            b.march(2)
        _log("  2 more steps, 48^3 slot 1 and 128^3 released")
        _same(grp, ref, "after release")
        # and back: a new load joins the next group step
        Q = _ic(128, 7)
        grp[128].load(0, Q, 150.0, nsteps=10)
        ref[128].load(0, Q, 150.0, nsteps=10)
        g.march(2)
        for b in ref.values():
            b.march(2)
        _log("  2 more steps, 128^3 reloaded")
        _same(grp, ref, "after reload")
    finally:
        g.close()
        _close(grp)
        _close(ref)


def test_group_slot_vs_oracle():
    """One 48^3 instance inside a group against the oracle, 3 steps."""
    n = 48
    re = st_amd.re_for_taylor(40.0, 4.0, 1.0, 0.3)
    bs = {m: st_amd.Batch(m, m, m, 1, **KW) for m in (48, 96)}
    Q = _ic(n, 12345)
    bs[48].load(0, Q, re, nsteps=3)
    bs[96].load(0, _ic(96, 5), 100.0, nsteps=3)
    g = st_amd.Group(bs.values())
    try:
        g.march(3)
        g.sync()
        Qg = bs[48].store(0)
        step, t, _, _ = bs[48].status()
        o = O.Oracle(n, n, n, Re=re, **KW)
        Qc, t_c, st_c, _ = o.march(Q, 3)
        e = relerr(Qg, Qc)
        _log("group 48^3 slot vs oracle, 3 steps:", e)
        assert step[0] == st_c == 3
        assert abs(t[0] - t_c) <= 1e-10 * t_c
        assert e.max() < 1e-10
    finally:
        g.close()
        _close(bs)


def test_group_rejects():
    b = st_amd.Batch(48, 48, 48, 1, **KW)
    c = st_amd.Batch(48, 48, 48, 1, **KW)
    d = st_amd.Batch(32, 32, 32, 1, **KW)
    e = st_amd.Batch(64, 64, 64, 1, Pr=0.7, hyper_param=0)
    try:
        with pytest.raises(st_amd.StError, match="one batch per size"):
            st_amd.Group([b, c])
        with pytest.raises(st_amd.StError, match="48\\^3, 64\\^3"):
            st_amd.Group([d])
        with pytest.raises(st_amd.StError, match="hyper_param"):
            st_amd.Group([b, e])
        g = st_amd.Group([b])           # nothing loaded: a march does nothing
        g.march(2)
        g.sync()
        g.close()
    finally:
        for x in (b, c, d, e):
            x.close()
