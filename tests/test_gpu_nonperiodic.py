"""GPU parity of the non-periodic x link set (src/makefile_nonperiodic1d:10-17).

The HIP path solves the closure-row systems as the recurrences of the
periodic band without the wrap plus a rank-4 (derivatives) or rank-6
(filter) boundary correction; the oracle runs pentadiagonal_serial
(src/pentadiagonal.f90:319-374).  Both are exact up to rounding, so the
differences scale with the condition numbers of the closure matrices
(about 15 for D1, 4e3 for D2 and 2e3 for the filter).  Tolerances,
max|gpu - cpu| / max|cpu| per field:
  D1, Gaussian                 1e-13
  D2, compact filter           1e-12
  one RHS / Hyper_RHS / stage  1e-12, or twice the oracle's own spread
                               under 1-ulp perturbations of Q where that is
                               larger
  artificial coefficients      1e-11, or three times that spread
  20 steps                     1e-10, or ten times the oracle's spread
                               (at most 1e-8)
The one-sided rows cancel strongly (the D2 row 1 sums |coefficients| 56
times its result, and |d4 e| passes two of them), so the reference's own
result moves by more than 1e-12 when Q moves by an ulp once the lines are
long: at 256 x-points the RHS spread is about 1e-11 and the kappa* spread
about 1e-9.  The one-sided rows and the primitive variables are rounded as
the reference rounds them (no contraction), which keeps the GPU well
inside that spread.
"""
import sys

import numpy as np
import pytest

import oracle as O
from fields import along, relerr, smooth_state

pytestmark = pytest.mark.gpu

st_amd = pytest.importorskip("st_amd")

PER = (2.0, 2 * np.pi, 2 * np.pi)

# x lines of 16..512 points: lanes per line P = 8, 16, 32, 64 and both
# per-lane widths of each pass
SHAPES = [(16, 32, 16), (32, 16, 16), (64, 16, 16), (128, 16, 32), (256, 16, 16), (512, 16, 16),
          (48, 16, 16), (192, 16, 16), (80, 16, 16), (160, 16, 16), (448, 16, 16)]


def _log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def solvers():
    cache = {}

    def get(shape, **kw):
        key = (shape, tuple(sorted(kw.items())))
        if key not in cache:
            nx, ny, nz = shape
            cache[key] = (st_amd.Solver(nx, ny, nz, period=PER, nonperiodic_x=1, **kw),
                          O.Oracle(nx, ny, nz, period=PER, nonper_x=1, **kw))
        return cache[key]
    yield get
    for s, _ in cache.values():
        s.close()


@pytest.mark.parametrize("shape", SHAPES)
def test_np_operators(solvers, shape):
    _log("np operators", shape)
    s, o = solvers(shape)
    rng = np.random.default_rng(3)
    F = np.asfortranarray(rng.normal(size=o.shape))
    G = np.asfortranarray(rng.normal(size=o.shape))
    for d in range(3):
        e = relerr(s.dfd(d, F), o.dfd(d, F))
        assert e < 1e-13, (d, e)
        e = relerr(s.dfd(d, F, dF=G, iadd=1, cadd=-0.5), o.dfd(d, F, dF=G, iadd=1, cadd=-0.5))
        assert e < 1e-13, (d, e)
        g1, g2 = s.d2fd2(d, F, iadd=1)
        c1, c2 = o.d2fd2(d, F, iadd=1)
        assert relerr(g1, c1) < 1e-13, d
        e = relerr(g2, c2)
        assert e < 1e-12, (d, e)
        e = relerr(s.filter_dir(d, F), o.filter_dir(d, F))
        assert e < 1e-12, (d, e)
        e = relerr(s.gauss_dir(d, F), o.gauss_dir(d, F))
        assert e < 1e-13, (d, e)


@pytest.mark.parametrize("shape", [(64, 16, 16), (256, 16, 16)])
def test_np_polynomial_exact(solvers, shape):
    """Closed form on the GPU alone: the closures make D1 exact for cubics
    and D2 for quartics; the filter keeps cubics.  The D2 rounding grows as
    n^2 times the closure matrix's condition (the oracle's own error is
    about 2.4e-15 n^2 here)."""
    s, _ = solvers(shape)
    n = shape[0]
    x = np.arange(n) * PER[0] / (n - 1) - 0.7
    F = along(0, 0.3 + x - 0.8 * x ** 2 + 0.25 * x ** 3, (shape[1], shape[2], n))
    d1 = along(0, 1 - 1.6 * x + 0.75 * x ** 2, F.shape)
    assert relerr(s.dfd(0, F), d1) < 1e-12
    F4 = along(0, 0.3 + x - 0.8 * x ** 2 + 0.25 * x ** 3 - 0.1 * x ** 4, F.shape)
    d2 = along(0, -1.6 + 1.5 * x - 1.2 * x ** 2, F.shape)
    assert relerr(s.d2fd2(0, F4), d2) < 2e-14 * n * n
    assert relerr(s.filter_dir(0, F), F) < 1e-12


@pytest.mark.parametrize("hyper", [0, 1])
@pytest.mark.parametrize("shape", [(32, 16, 16), (128, 16, 32), (256, 16, 16), (96, 48, 16)])
def test_np_rhs(solvers, shape, hyper):
    _log("np rhs", shape, "hyper", hyper)
    s, o = solvers(shape, Re=80.0, hyper_param=hyper)
    Q = smooth_state(*shape, seed=21)
    # the oracle's spread under 1-ulp perturbations of Q
    Qp = np.asfortranarray(Q * (1 + np.random.default_rng(0).choice([-1, 1], size=Q.shape) * 2.2e-16))
    rp, _ = o.rhs(Qp)
    hp = o.hyper_rhs(Qp, rp)
    Hp = o.hyper().copy() if hyper else None
    r_gpu = s.rhs(Q)
    r_cpu, _ = o.rhs(Q)
    e = relerr(r_gpu, r_cpu)
    _log("  NS rhs relerr", e)
    h_gpu = s.hyper_rhs(r_gpu)
    h_cpu = o.hyper_rhs(Q, r_cpu)
    spread = relerr(hp, h_cpu).max()
    tol = max(1e-12, 2 * spread)
    _log("  ulp spread of the oracle", spread)
    assert e.max() < tol
    e = relerr(h_gpu, h_cpu)
    _log("  full rhs relerr", e)
    assert e.max() < tol
    if hyper:
        eh = relerr(s.hyper(), o.hyper())
        sh = relerr(Hp, o.hyper())
        _log("  hyper coeff relerr", eh, "spread", sh)
        assert (eh < np.maximum(1e-11, 3 * sh)).all()


@pytest.mark.parametrize("shape", [(64, 16, 16), (256, 16, 16)])
def test_np_stage(solvers, shape):
    s, o = solvers(shape, Re=80.0, hyper_param=1)
    Q = smooth_state(*shape, seed=8)
    QW = np.zeros_like(Q)
    s.upload(Q)
    s.rk_stage(0.0, 0.1028639988104959, 1e-3, reset_qw=True)
    Qc, _, _ = o.rk_stage(Q, QW, 0.0, 0.1028639988104959, 1e-3)
    e = relerr(s.download(), Qc)
    _log("np stage relerr", e)
    assert e.max() < 1e-12


def test_np_steps_20(solvers):
    shape = (64, 16, 16)
    s, o = solvers(shape, Re=200.0, hyper_param=1)
    Q = smooth_state(*shape, seed=5, amp=0.3)
    s.upload(Q)
    t_g, st_g, dts_g = s.march(20)
    Qg = s.download()
    Qc, t_c, st_c, dts_c = o.march(Q, 20)
    # the oracle's own spread after 20 steps from a 1-ulp perturbed start
    Qp = np.asfortranarray(Q * (1 + np.random.default_rng(0).choice([-1, 1], size=Q.shape) * 2.2e-16))
    o2 = O.Oracle(*shape, period=PER, nonper_x=1, Re=200.0, hyper_param=1)
    Qc2, _, _, _ = o2.march(Qp, 20)
    spread = relerr(Qc2, Qc).max()
    e = relerr(Qg, Qc)
    _log("np 20 steps relerr", e, "spread", spread, "dt", dts_g[0], dts_c[0])
    assert st_g == st_c == 20
    assert abs(t_g - t_c) <= 1e-10 * t_c
    # the GPU differs from the oracle by a few ulps at every operation, not
    # only at the start, so its distance after 20 steps is a small multiple
    # of the single-perturbation spread
    assert e.max() < min(1e-8, max(1e-10, 10 * spread))


def test_np_shock_sponge_steps():
    """Config-2 pieces together: a Mach-1.5 shock with upstream turbulence,
    Rankine-Hugoniot sponge layers at both ends, 10 steps on the GPU against
    the oracle (tolerance from the oracle's own 1-ulp spread)."""
    shape = (64, 16, 16)
    Q, s1, s2 = st_amd.ic_shock(*shape, PER, mach=1.5, frac=0.25, urms=0.1, seed=3)
    kw = dict(Re=300.0, hyper_param=1)
    s = st_amd.Solver(*shape, period=PER, nonperiodic_x=1, **kw)
    try:
        s.sponge(5.0, 2.0, 0.4, s1, s2)
        s.upload(Q)
        t_g, st_g, _ = s.march(10)
        Qg = s.download()
    finally:
        s.close()
    outs = []
    for pert in (0, 1):
        o = O.Oracle(*shape, period=PER, nonper_x=1, **kw)
        o.set_sponge(5.0, 2.0, 0.4, s1, s2)
        Qi = Q if not pert else np.asfortranarray(
            Q * (1 + np.random.default_rng(0).choice([-1, 1], size=Q.shape) * 2.2e-16))
        outs.append(o.march(Qi, 10))
    Qc, t_c = outs[0][0], outs[0][1]
    spread = relerr(outs[1][0], Qc).max()
    e = relerr(Qg, Qc)
    _log("shock+sponge 10 steps relerr", e, "spread", spread)
    assert st_g == 10 and abs(t_g - t_c) <= 1e-10 * t_c
    assert e.max() < min(1e-8, max(1e-10, 10 * spread))
    assert np.isfinite(Qg).all()
