"""GPU parity of the start-up path of Program Main (src/main.f:96-104):
Compute_Grad (src/grad.f:13-115, the log's MaxDil and Enstrophy),
Compute_hyper before the first time step (src/hyper_dynamic.f:173-318;
non-periodic x: src/hyper_routines.f:66-214), and the dt limits it feeds
(src/lddrk4_periodic_sf.f:99-142), through the C-ABI against the oracle.

Tolerances, max|gpu - cpu| / max|cpu| per field: gradients 1e-13,
coefficients 1e-11 (1e-10 where a non-periodic x doubles the one-sided
D2 rows, whose spread under 1-ulp moves of Q is of that size), MaxDil and
Enstrophy 1e-12 relative (reduction order differs), steps 1e-10.
"""
import sys

import numpy as np
import pytest

import oracle as O
from fields import relerr, smooth_state

pytestmark = pytest.mark.gpu

st_amd = pytest.importorskip("st_amd")

PER_NP = (2.0, 2 * np.pi, 2 * np.pi)


def _log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("nonper", [0, 1])
@pytest.mark.parametrize("shape", [(32, 32, 32), (64, 32, 16)])
def test_compute_hyper_fresh_and_after_rhs(shape, nonper):
    _log("compute_hyper", shape, "nonper", nonper)
    per = PER_NP if nonper else (2 * np.pi,) * 3
    kw = dict(Re=80.0, Pr=0.7, hyper_param=1)
    Q = smooth_state(*shape, seed=21, amp=1.2, mt=0.5)
    if nonper:
        Q, _, _ = st_amd.ic_shock(*shape, per, mach=1.5, frac=0.25, urms=0.1, seed=3)
    s = st_amd.Solver(*shape, period=per, nonperiodic_x=nonper, **kw)
    o = O.Oracle(*shape, period=per, nonper_x=nonper, **kw)
    tol = 1e-10 if nonper else 1e-11
    try:
        # a new context: /GRAD/ is zero, so mu* = beta* = 0 and kappa* comes from e
        Hg = s.compute_hyper(Q)
        Hc = o.compute_hyper(Q)
        assert np.all(Hc[..., 1] == 0.0) and np.all(Hg[..., 1] == 0.0)
        assert np.abs(Hc[..., 2]).max() > 0
        e = relerr(Hg, Hc)
        _log("  fresh", e)
        assert e.max() < tol
        # after an RHS of the same state: the gradients it left
        s.rhs(Q)
        o.rhs(Q)
        e = relerr(s.compute_hyper(), o.compute_hyper(Q))
        _log("  after rhs", e)
        assert e.max() < tol
    finally:
        s.close()


@pytest.mark.parametrize("shape", [(32, 32, 32), (48, 32, 64)])
def test_compute_grad(shape):
    _log("compute_grad", shape)
    s = st_amd.Solver(*shape, Re=80.0, hyper_param=1)
    o = O.Oracle(*shape, Re=80.0, hyper_param=1)
    try:
        Q = st_amd.ic_isotropic(*shape, k0=8.0, urms=1.0, mt=0.5, seed=8)
        s.upload(Q)
        mg, eg = s.compute_grad()
        mc, ec = o.compute_grad(Q)
        _log("  MaxDil", mg, mc, "Enstrophy", eg, ec)
        assert abs(mg - mc) <= 1e-12 * mc and abs(eg - ec) <= 1e-12 * ec
        assert relerr(s.grad()[..., :9], o.grad()[..., :9]).max() < 1e-13
    finally:
        s.close()


def test_resume_with_binding_hyper_limit():
    """A run resumed at step 3 in the shocklet regime (k0 = 8, M_t = 0.5):
    Compute_Grad (out_init = 1) and Compute_hyper at start-up make the
    artificial-transport limit bind on the first resumed step, on both
    sides, and the steps that follow agree with the oracle."""
    n = 32
    k0, mt = 8.0, 0.5
    re = st_amd.re_for_taylor(40.0, k0, 1.0, mt)
    kw = dict(Re=re, Pr=0.7, hyper_param=1)
    o0 = O.Oracle(n, n, n, **kw)
    Q0 = st_amd.ic_isotropic(n, n, n, k0=k0, urms=1.0, mt=mt, seed=5)
    Q3, t3, s3, _ = o0.march(Q0, 3)
    assert s3 == 3
    o = O.Oracle(n, n, n, **kw)
    o.compute_grad(Q3)
    o.compute_hyper(Q3)
    _, rv = o.cfl(Q3)
    dt_cfl = 0.5 / rv
    dt_first = o.tstep(Q3, rv, s3)
    assert dt_first < 0.9 * dt_cfl                       # the limit binds
    Qc, t_c, st_c, dts_c = o.march(Q3, 5, time=t3, step=s3)
    s = st_amd.Solver(n, n, n, **kw)
    try:
        s.upload(Q3)
        s.compute_grad()
        s.compute_hyper()
        t_g, st_g, dts_g = s.march(5, time=t3, step=s3)
        Qg = s.download()
    finally:
        s.close()
    _log("resume: dt", dts_g[0], dts_c[0], "cfl dt", dt_cfl)
    assert abs(dts_g[0] - dts_c[0]) <= 1e-12 * dts_c[0]
    assert np.all(np.abs(np.array(dts_g) - np.array(dts_c)) <= 1e-10 * dts_c[0])
    # the time is the sum of the dts, each within the bound above
    assert st_g == st_c == s3 + 5 and abs(t_g - t_c) <= 1e-10 * t_c
    assert relerr(Qg, Qc).max() < 1e-10


def test_nonperiodic_first_step_limit():
    """Config 2's rule applies the limits from the first step
    (src/lddrk4_nonperiodic_sf.f:105) with the coefficients of the start-up
    Compute_hyper: kappa* of the shock's energy profile makes that limit
    finite at step 1 (it does not bind for this smooth shock)."""
    shape = (64, 16, 16)
    Q, s1, s2 = st_amd.ic_shock(*shape, PER_NP, mach=1.5, frac=0.25, urms=0.1, seed=3)
    kw = dict(Re=300.0, hyper_param=1)
    o = O.Oracle(*shape, period=PER_NP, nonper_x=1, **kw)
    o.set_sponge(5.0, 2.0, 0.4, s1, s2)
    H = o.compute_hyper(Q)
    assert H[..., 2].max() > 0
    _, rv = o.cfl(Q)
    Qc, t_c, st_c, dts_c = o.march(Q, 3)
    s = st_amd.Solver(*shape, period=PER_NP, nonperiodic_x=1, **kw)
    try:
        s.sponge(5.0, 2.0, 0.4, s1, s2)
        s.upload(Q)
        s.compute_hyper()
        t_g, st_g, dts_g = s.march(3)
        Qg = s.download()
    finally:
        s.close()
    _log("np first-step dt", dts_g[0], dts_c[0], "cfl", 0.5 / rv)
    assert abs(dts_g[0] - dts_c[0]) <= 1e-10 * dts_c[0]
    assert relerr(Qg, Qc).max() < 1e-9
