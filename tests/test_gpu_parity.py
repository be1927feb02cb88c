"""GPU parity: the HIP path (through the C-ABI) against the CPU oracle.

Tolerances are norm-wise relative per field, max|gpu - cpu| / max|cpu|:
  compact D1, D2, Gaussian         1e-13
  compact filter (cond ~ 2e3)      1e-12
  one RHS / Hyper_RHS / one stage  1e-12
  hyper coefficients               1e-11 (x (n/256)^4 for lines n > 256)
  20 steps                         1e-10
  200 steps                        1e-9   (the north-star bar)
The GPU differs from the oracle only by rounding: fused multiply-adds,
D1 of a sum instead of a sum of D1s, and the cyclic solves done as
factored recurrences instead of bordered elimination.
"""
import sys

import numpy as np
import pytest

import oracle as O
from fields import along, grid_1d, relerr, smooth_state

pytestmark = pytest.mark.gpu

st_amd = pytest.importorskip("st_amd")


def _log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


# Line lengths 16..256 cover every lane layout of the passes: lines of
# P = 8, 16, 32 and 64 lanes (one to eight lines per wave).  Lines of 48,
# 96, 192 and 384 points use 6 (or 3) points per lane, lines of 40..320
# five and of 56..448 seven.
SHAPES = [(32, 32, 32), (64, 32, 16), (16, 64, 32), (256, 16, 16), (16, 128, 32), (16, 16, 256),
          (48, 96, 16), (16, 48, 192), (384, 16, 16)]
LONG = [(256, 16, 16), (16, 128, 32), (32, 16, 256)]
THREE = [(48, 48, 48), (96, 16, 48), (16, 192, 16)]
# lines of five and seven times a power of two (5 or 7 points per lane)
FIVE_SEVEN = [(80, 16, 16), (16, 112, 16), (16, 16, 160), (40, 56, 16), (320, 16, 16), (16, 448, 16),
              (16, 16, 224)]


@pytest.fixture(scope="module")
def solvers():
    cache = {}

    def get(shape, **kw):
        key = (shape, tuple(sorted(kw.items())))
        if key not in cache:
            nx, ny, nz = shape
            cache[key] = (st_amd.Solver(nx, ny, nz, **kw), O.Oracle(nx, ny, nz, **kw))
        return cache[key]
    yield get
    for s, _ in cache.values():
        s.close()


@pytest.mark.parametrize("shape", SHAPES + FIVE_SEVEN)
def test_operators(solvers, shape):
    _log("operators", shape)
    s, o = solvers(shape)
    rng = np.random.default_rng(1)
    F = np.asfortranarray(rng.normal(size=o.shape))
    G = np.asfortranarray(rng.normal(size=o.shape))
    for d in range(3):
        assert relerr(s.dfd(d, F), o.dfd(d, F)) < 1e-13
        assert relerr(s.dfd(d, F, dF=G, iadd=1, cadd=-0.5), o.dfd(d, F, dF=G, iadd=1, cadd=-0.5)) < 1e-13
        g1, g2 = s.d2fd2(d, F, iadd=1)
        c1, c2 = o.d2fd2(d, F, iadd=1)
        assert relerr(g1, c1) < 1e-13
        assert relerr(g2, c2) < 1e-13
        assert relerr(s.filter_dir(d, F), o.filter_dir(d, F)) < 1e-12
        assert relerr(s.gauss_dir(d, F), o.gauss_dir(d, F)) < 1e-13
    Q = np.asfortranarray(rng.normal(size=o.shape + (2,)))
    assert relerr(s.compact_filter(Q), o.compact_filter(Q)).max() < 1e-12
    assert relerr(s.gaussian_filter(Q), o.gaussian_filter(Q)).max() < 1e-13


@pytest.mark.parametrize("shape", [(32, 32, 32), (64, 32, 16)])
def test_d1_sinusoid(solvers, shape):
    """Closed-form check straight on the GPU (independent of the oracle)."""
    s, o = solvers(shape)
    from fields import sym_d1
    for d in range(3):
        n = shape[d]
        h = 2 * np.pi / n
        x = grid_1d(n)
        for k in (1, 3, n // 2 - 1):
            F = along(d, np.sin(k * x), o.shape)
            ex = along(d, sym_d1(k * h) / h * np.cos(k * x), o.shape)
            assert np.abs(s.dfd(d, F) - ex).max() < 1e-13 * max(1, k)


# x lines of 512 and 448 points: the default launch shapes of some passes
# (1024-thread tiles) do not fit them, so the passes fall back to a tile the
# grid takes
WIDE_X = [(512, 16, 16), (448, 16, 16)]


@pytest.mark.parametrize("hyper", [0, 1])
@pytest.mark.parametrize("shape", [(32, 32, 32), (64, 32, 16)] + LONG + THREE + [(40, 56, 112)] + WIDE_X)
def test_rhs(solvers, shape, hyper):
    _log("rhs", shape, "hyper", hyper)
    s, o = solvers(shape, Re=80.0, hyper_param=hyper)
    Q = smooth_state(*shape, seed=11)
    r_gpu = s.rhs(Q)
    r_cpu, _ = o.rhs(Q)
    e = relerr(r_gpu, r_cpu)
    _log("  NS rhs relerr", e)
    assert e.max() < 1e-12
    assert relerr(s.grad(), o.grad()).max() < 1e-12
    h_gpu = s.hyper_rhs(r_gpu)
    h_cpu = o.hyper_rhs(Q, r_cpu)
    e = relerr(h_gpu, h_cpu)
    _log("  full rhs relerr", e)
    assert e.max() < 1e-12
    if hyper:
        # the coefficients come from fourth derivatives (Lap Lap |S|, Lap Lap
        # e), whose rounding grows as h^-4: the bar scales with the longest
        # line (the right-hand side built from them agrees to 1e-12 above)
        eh = relerr(s.hyper(), o.hyper())
        _log("  hyper coeff relerr", eh)
        assert eh.max() < 1e-11 * max(1.0, (max(shape) / 256.0) ** 4)


def test_rhs_turbulence_ic(solvers):
    shape = (32, 32, 32)
    s, o = solvers(shape, Re=300.0, hyper_param=1)
    Q = st_amd.ic_isotropic(*shape, k0=4.0, urms=1.0, mt=0.3, seed=7)
    e = relerr(s.full_rhs(Q), o.full_rhs(Q))
    _log("ic rhs relerr", e)
    assert e.max() < 1e-12


def test_reductions_and_fix(solvers):
    shape = (32, 32, 32)
    s, o = solvers(shape, Re=80.0, hyper_param=1)
    Q = smooth_state(*shape, seed=2)
    s.upload(Q)
    st, rv = s.cfl()
    ost, orv = o.cfl(Q)
    assert abs(st - ost) <= 1e-13 * ost and abs(rv - orv) <= 1e-13 * orv
    assert abs(s.tstep(rv, 0) - o.tstep(Q, orv, 0)) <= 1e-14 * o.tstep(Q, orv, 0)
    H = np.asfortranarray(np.abs(np.random.default_rng(3).normal(size=o.shape + (3,))) * 1e-3 + 1e-6)
    s.set_hyper(H)
    o.set_hyper(H)
    assert abs(s.tstep(rv, 5) - o.tstep(Q, orv, 5)) <= 1e-13 * o.tstep(Q, orv, 5)
    np.testing.assert_allclose(s.diagnostics(), o.diagnostics(Q), rtol=1e-12, atol=1e-14)
    Qn = Q.copy()
    Qn[3, 4, 5, 4] = 0.0
    Qn[0, 1, 2, 4] = -1.0
    assert relerr(s.fix_pressure(Qn), o.fix_pressure(Qn)).max() < 1e-15


@pytest.mark.parametrize("hyper", [0, 1])
@pytest.mark.parametrize("shape", [(32, 32, 32)] + LONG + THREE[:2] + WIDE_X[:1])
def test_stage(solvers, shape, hyper):
    _log("stage", shape, "hyper", hyper)
    s, o = solvers(shape, Re=80.0, hyper_param=hyper)
    Q = smooth_state(*shape, seed=4)
    QW = np.zeros_like(Q)
    s.upload(Q)
    s.rk_stage(0.0, 0.1028639988104959, 2e-3, reset_qw=True)
    Qc, QWc, _ = o.rk_stage(Q, QW, 0.0, 0.1028639988104959, 2e-3)
    e = relerr(s.download(), Qc)
    _log("stage relerr", e)
    assert e.max() < 1e-12


def _run_steps(solvers, shape, nsteps, tol, **kw):
    s, o = solvers(shape, **kw)
    Q = st_amd.ic_isotropic(*shape, k0=4.0, urms=1.0, mt=0.3, seed=12345)
    s.upload(Q)
    t_g, st_g, dts_g = s.march(nsteps)
    Qg = s.download()
    Qc, t_c, st_c, dts_c = o.march(Q, nsteps)
    e = relerr(Qg, Qc)
    _log("%d steps relerr" % nsteps, e, "dt rel", np.abs(np.array(dts_g) - np.array(dts_c)).max() / dts_c[0])
    assert st_g == st_c == nsteps
    assert abs(t_g - t_c) <= 1e-10 * t_c
    assert e.max() < tol


def test_steps_20(solvers):
    re = st_amd.re_for_taylor(40.0, 4.0, 1.0, 0.3)
    _run_steps(solvers, (32, 32, 32), 20, 1e-10, Re=re, Pr=0.7, hyper_param=1)


def test_steps_20_48(solvers):
    """The smallest served and batched grid (48^3: 6 points per lane)."""
    re = st_amd.re_for_taylor(40.0, 4.0, 1.0, 0.3)
    _run_steps(solvers, (48, 48, 48), 20, 1e-10, Re=re, Pr=0.7, hyper_param=1)


def test_steps_20_80():
    """80^3 (5 points per lane in every direction), 20 steps."""
    re = st_amd.re_for_taylor(40.0, 4.0, 1.0, 0.3)
    cache = {}

    def get(shape, **kw):
        if shape not in cache:
            cache[shape] = (st_amd.Solver(*shape, **kw), O.Oracle(*shape, **kw))
        return cache[shape]
    try:
        _run_steps(get, (80, 80, 80), 20, 1e-10, Re=re, Pr=0.7, hyper_param=1)
    finally:
        for s, _ in cache.values():
            s.close()


@pytest.mark.slow
def test_steps_200(solvers):
    re = st_amd.re_for_taylor(40.0, 4.0, 1.0, 0.3)
    _run_steps(solvers, (32, 32, 32), 200, 1e-9, Re=re, Pr=0.7, hyper_param=1)
