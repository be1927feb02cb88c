"""Serving and ensembles over batches: every run's result is the same bits
alone (plain path, st_march) and in the mix, whatever else is running, and
admission respects the memory budget."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

st_amd = pytest.importorskip("st_amd")
import st_serve  # noqa: E402


def _log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_served_runs_equal_alone():
    runs = st_serve.poisson_trace(8, rate=200.0, sizes=(32, 48), steps=(2, 6), seed=7)
    # room for two slots per size: later runs wait for a slot
    budget = 2 * st_serve.slot_bytes(32) + 2 * st_serve.slot_bytes(48)
    srv = st_serve.Server(runs, budget_bytes=budget, slice_s=0.002)
    try:
        assert sum(st_serve.slot_bytes(n) * k for n, k in srv.slots.items()) <= budget
        assert sum(srv.slots.values()) < len(runs)       # some runs must wait for a slot
        st = srv.serve()
    finally:
        srv.close()
    _log("served", st)
    plain, ref = st_serve.plain_one_at_a_time(runs)
    _log("plain", plain)
    for r in runs:
        assert r.done >= r.admit >= 0.0
        assert np.array_equal(r.result, ref[r.rid]), r.rid
    assert st["runs"] == 8 and st["latency_p99_s"] >= st["latency_p50_s"] > 0


def test_ensemble_equal_alone():
    runs = [st_serve.Run(i, n, 3, seed=i + 1, k0=(3.0, 4.0, 6.0)[i % 3]) for i, n in enumerate([32, 32, 48, 48, 48])]
    ens = st_serve.Ensemble(runs)
    try:
        ens.march(3)
        got = ens.results()
    finally:
        ens.close()
    _, ref = st_serve.plain_one_at_a_time(runs)
    for r in runs:
        assert np.array_equal(got[r.rid], ref[r.rid]), r.rid


def test_autotune_keeps_results():
    """Tuning restores the state bit for bit.  The tile width does not
    change results; the lane layout does at rounding level (the recurrences
    are chunked M points per lane), so the tuned march is compared with
    the untuned one at the parity tolerance."""
    shape = (48, 48, 48)
    Q = st_amd.ic_isotropic(*shape, k0=4.0, urms=1.0, mt=0.3, seed=9)
    b = st_amd.Batch(*shape, 2, Re=200.0)
    try:
        b.load(0, Q, 200.0, 3)
        cfg = b.autotune(1)
        _log("tuned", {k: v for k, v in cfg.items() if k.split(".")[0] in ("grad", "hyper_div", "hdiv_rkf")})
        assert np.array_equal(b.store(0), Q)
        b.march(3)
        b.sync()
        s = st_amd.Solver(*shape, Re=200.0)
        try:
            s.upload(Q)
            s.march(3)
            from fields import relerr
            assert relerr(b.store(0), s.download()).max() < 1e-11    # 3 steps (33 stages)
        finally:
            s.close()
    finally:
        b.close()


def test_tuned_batches_and_plain_path_share_bits():
    """Batches that tune their launch shapes hand them to the plain path
    (st_pass_config_set), so every run still gives the same bits alone and
    in the mix."""
    runs = [st_serve.Run(i, n, 3, seed=i + 1, k0=(3.0, 4.0, 6.0)[i % 3]) for i, n in enumerate([32, 32, 48, 48])]
    ens = st_serve.Ensemble(runs, autotune=1)
    try:
        ens.march(3)
        got = ens.results()
        tuned = dict(ens.tuned)
    finally:
        ens.close()
    _log("tuned 48", {k: v for k, v in tuned[48].items() if k.split(".")[0] in ("grad", "grad_cross", "hdiv_rkf")})
    _, ref = st_serve.plain_one_at_a_time(runs, configs=tuned)
    for r in runs:
        assert np.array_equal(got[r.rid], ref[r.rid]), r.rid
    s = st_amd.Solver(48, 48, 48, Re=200.0)
    try:
        s.set_pass_configs(tuned[48])
        assert s.pass_configs() == tuned[48]
    finally:
        s.close()
