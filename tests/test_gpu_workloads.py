"""The benchmark workloads themselves, on the GPU through the C-ABI, against
the oracle where the oracle finishes in about a minute and through
size-independent properties where it does not.

  64^3, 200 steps (k0 = 4, M_t = 0.3 and the shocklet regime k0 = 8,
      M_t = 0.5): the march against the oracle's march, 1e-9 norm-wise
      relative; in the shocklet regime the dilatation sensor, the
      artificial-transport dt limit and the pressure clamp are active
      (asserted on the oracle's run).
  256^3 (k0 = 8, M_t = 0.5): one full step (11 stages) against the oracle,
      1e-11; then 200 steps on the GPU: finite, mass conserved to rounding,
      mean momentum stays at the level of the scheme's non-conservative
      convective form, time increasing, dt positive.
  512x192x192 (non-periodic x, Mach 1.5 shock, sponge): one step against
      the oracle (1e-10, the one-sided rows' 1-ulp spread at 512 points),
      then 500 steps on the GPU: finite, and the sponge cores hold the
      Rankine-Hugoniot states.
  The 48-instance ensemble (16 x 48^3, 16 x 64^3, 12 x 96^3, 4 x 128^3),
      100 steps: every slot bit-identical to its run alone on the plain
      path; one slot per size against the oracle over 5 steps, 1e-10.
  The served 16-run trace (Poisson arrivals, N in {48..128}^3, 50-400
      steps): every run's result bit-identical alone and in the mix; the
      plain path per size against the oracle over 5 steps, 1e-10.
"""
import os
import sys
import time

import numpy as np
import pytest

import oracle as O
from fields import relerr

pytestmark = pytest.mark.gpu

st_amd = pytest.importorskip("st_amd")

O.lib().or_threads(min(16, os.cpu_count() or 1))


def _log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _oracle_march(o, Q, nsteps, time0=0.0, step0=0):
    """src/main.f:129-136 on the oracle, also counting the steps whose dt the
    artificial-transport limit set (dt < cfl_cal / cfl_resv)."""
    t, step, bound, dts = time0, step0, 0, []
    for _ in range(nsteps):
        _, rv = o.cfl(Q)
        dt = o.tstep(Q, rv, step)
        bound += dt < o.params.cfl_cal / rv
        Q, t, step = o.lddrk4(Q, dt, time=t, step=step)
        dts.append(dt)
    return Q, t, step, dts, bound


@pytest.mark.parametrize("k0,mt", [(4.0, 0.3), (8.0, 0.5)])
def test_64cubed_200_steps(k0, mt):
    n, nsteps = 64, 200
    re = st_amd.re_for_taylor(40.0, k0, 1.0, mt)
    kw = dict(Re=re, Pr=0.7, hyper_param=1)
    Q0 = st_amd.ic_isotropic(n, n, n, k0=k0, urms=1.0, mt=mt, seed=12345)
    _log("64^3 200 steps k0 %g Mt %g: GPU" % (k0, mt))
    s = st_amd.Solver(n, n, n, **kw)
    try:
        s.upload(Q0)
        t_g, st_g, dts_g = s.march(nsteps)
        Qg = s.download()
    finally:
        s.close()
    _log("  oracle")
    o = O.Oracle(n, n, n, **kw)
    t0 = time.time()
    Qc, t_c, st_c, dts_c, bound = _oracle_march(o, Q0, nsteps)
    H = o.hyper()
    cb_max = (2.0e-3 * H[..., 0] / H[..., 1]).max()
    e = relerr(Qg, Qc)
    _log("  relerr", e, "oracle %.1f s" % (time.time() - t0), "limit-bound steps", bound,
         "max C_beta %.3f" % cb_max, "clamped points", o.clamps())
    assert st_g == st_c == nsteps and abs(t_g - t_c) <= 1e-10 * t_c
    assert np.abs(np.array(dts_g) - np.array(dts_c)).max() <= 1e-9 * min(dts_c)
    assert e.max() < 1e-9
    if k0 == 8.0:
        # the shocklet regime: the sensor switches on, the artificial
        # transport sets the time step, and Fix_Pressure clamps negative
        # pressures (4 points over these 200 steps)
        assert bound > 0 and cb_max > 0.5 and o.clamps() > 0


def test_256cubed_step_and_200_steps():
    n = 256
    k0, mt = 8.0, 0.5
    re = st_amd.re_for_taylor(40.0, k0, 1.0, mt)
    kw = dict(Re=re, Pr=0.7, hyper_param=1)
    Q0 = st_amd.ic_isotropic(n, n, n, k0=k0, urms=1.0, mt=mt, seed=12345)
    s = st_amd.Solver(n, n, n, **kw)
    try:
        _log("256^3: one step on the GPU")
        s.upload(Q0)
        t_g, st_g, dts_g = s.march(1)
        Q1 = s.download()
        _log("  oracle step")
        o = O.Oracle(n, n, n, **kw)
        t0 = time.time()
        Qc, t_c, st_c, dts_c = o.march(Q0, 1)
        e = relerr(Q1, Qc)
        _log("  relerr", e, "oracle %.1f s" % (time.time() - t0))
        del o
        assert abs(dts_g[0] - dts_c[0]) <= 1e-13 * dts_c[0]
        assert e.max() < 1e-11
        _log("  199 more steps on the GPU")
        d0 = s.diagnostics()
        t, step, dts = t_g, st_g, [dts_g[0]]
        for _ in range(199 // 20 + 1):
            k = min(20, 200 - step)
            if k <= 0:
                break
            t1, step, dk = s.march(k, time=t, step=step)
            assert t1 > t
            t = t1
            dts.extend(dk)
            _log("    step", step)
        Qe = s.download()
        d1 = s.diagnostics()
    finally:
        s.close()
    assert step == 200 and np.isfinite(Qe).all()
    assert np.all(np.array(dts) > 0) and np.all(np.isfinite(dts))
    _log("  diagnostics start", d0, "end", d1)
    assert abs(d1[4] - d0[4]) <= 1e-12 * d0[4]                 # mass
    assert np.abs(d1[:3]).max() < 1e-3                           # mean momentum (u_rms = 1)
    assert d1[3] < d0[3]                                         # decaying kinetic energy
    assert abs(d1[5] - d0[5]) < 1e-2 * d0[5]                     # total energy


def test_512x192x192_step_and_500_steps():
    shape = (512, 192, 192)
    # equal spacing in x, y, z: x spans 511 spacings of 2 pi / 192 (both end points)
    per = (2.0 * np.pi / 192 * 511, 2.0 * np.pi, 2.0 * np.pi)
    Q0, s1, s2 = st_amd.ic_shock(*shape, per, mach=1.5, frac=0.25, k0=4.0, urms=0.1, seed=12345)
    kw = dict(Re=500.0, Pr=0.7, hyper_param=1)
    s = st_amd.Solver(*shape, period=per, nonperiodic_x=1, **kw)
    try:
        s.sponge(5.0, 2.0, 0.4, s1, s2)
        _log("512x192x192: one step on the GPU")
        s.upload(Q0)
        s.compute_hyper()
        t_g, st_g, dts_g = s.march(1)
        Q1 = s.download()
        _log("  oracle step")
        o = O.Oracle(*shape, period=per, nonper_x=1, **kw)
        o.set_sponge(5.0, 2.0, 0.4, s1, s2)
        o.compute_hyper(Q0)
        t0 = time.time()
        Qc, t_c, st_c, dts_c = o.march(Q0, 1)
        e = relerr(Q1, Qc)
        _log("  relerr", e, "oracle %.1f s" % (time.time() - t0))
        del o
        assert abs(dts_g[0] - dts_c[0]) <= 1e-12 * dts_c[0]
        assert e.max() < 1e-10
        _log("  499 more steps on the GPU")
        t, step = t_g, st_g
        while step < 500:
            t1, step, dk = s.march(min(50, 500 - step), time=t, step=step)
            assert t1 > t and np.all(np.array(dk) > 0)
            t = t1
            _log("    step", step)
        Qe = s.download()
    finally:
        s.close()
    assert np.isfinite(Qe).all()
    # the sponge cores (the outer half of each layer) hold their states:
    # plane means of the conserved variables within 2% of the
    # Rankine-Hugoniot values (the turbulence is 0.1 of the inflow speed)
    nx = shape[0]
    lay = int(np.ceil(0.4 * nx / 2))
    for sl, ref in ((slice(0, lay // 2), s1), (slice(nx - lay // 2, nx), s2)):
        m = Qe[:, :, sl, :].mean(axis=(0, 1, 2))
        for l in (0, 3, 4):
            assert abs(m[l] - ref[l]) <= 0.02 * abs(ref[l]), (l, m[l], ref[l])


def _ensemble_slot_vs_oracle(runs, nsteps):
    """One run per size: the batched ensemble's slot after nsteps against the
    oracle's march of the same run."""
    import st_serve
    ens = st_serve.Ensemble(runs)
    try:
        ens.march(nsteps)
        got = ens.results()
    finally:
        ens.close()
    seen = set()
    for r in runs:
        if r.n in seen:
            continue
        seen.add(r.n)
        o = O.Oracle(r.n, r.n, r.n, Re=r.Re, Pr=0.7, hyper_param=1, cfl_cal=0.5)
        Qc, _, _, _ = o.march(r.make_ic(), nsteps)
        e = relerr(got[r.rid], Qc)
        _log("  slot of size %d vs oracle" % r.n, e)
        assert e.max() < 1e-10


def test_ensemble_48_instances():
    import st_serve
    _log("ensemble: 48 instances, 5 steps, one slot per size vs the oracle")
    _ensemble_slot_vs_oracle(st_serve.ensemble_config3(5), 5)
    _log("ensemble: 48 instances, 100 steps, each slot vs its run alone")
    runs = st_serve.ensemble_config3(100)
    ens = st_serve.Ensemble(runs)
    try:
        ens.march(100)
        got = ens.results()
        status = {n: p.batch.status() for n, p in ens.pools.items()}
    finally:
        ens.close()
    for n, (step, _, _, _) in status.items():
        assert np.all(step == 100), (n, step)
    _, alone = st_serve.plain_one_at_a_time(runs)
    for r in runs:
        assert np.array_equal(got[r.rid], alone[r.rid]), r.rid


def test_served_trace_16_runs():
    import st_serve
    runs = st_serve.poisson_trace(16, 20.0)
    _log("served trace: 16 runs, sizes", sorted(r.n for r in runs))
    srv = st_serve.Server(runs)
    try:
        st = srv.serve()
    finally:
        srv.close()
    _log("  served", st)
    _, alone = st_serve.plain_one_at_a_time(runs)
    for r in runs:
        assert r.result is not None and np.array_equal(r.result, alone[r.rid]), r.rid
    # the plain path of each size against the oracle over 5 steps
    seen = set()
    for r in runs:
        if r.n in seen:
            continue
        seen.add(r.n)
        s = st_amd.Solver(r.n, r.n, r.n, Re=r.Re, Pr=0.7, hyper_param=1, cfl_cal=0.5)
        try:
            s.upload(r.make_ic())
            s.march(5)
            Qg = s.download()
        finally:
            s.close()
        o = O.Oracle(r.n, r.n, r.n, Re=r.Re, Pr=0.7, hyper_param=1, cfl_cal=0.5)
        Qc, _, _, _ = o.march(r.make_ic(), 5)
        e = relerr(Qg, Qc)
        _log("  size %d plain path vs oracle, 5 steps" % r.n, e)
        assert e.max() < 1e-10
