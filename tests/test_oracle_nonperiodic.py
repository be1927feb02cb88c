"""The oracle's non-periodic x operators, pinned against mathematics.

The non-periodic link set (src/makefile_nonperiodic1d:10-17) closes the x
derivatives and filters with boundary rows (src/derivCDopt6_nonperiodic1d.f,
src/filters_nonperiodic1d.f) solved by pentadiagonal_serial
(src/pentadiagonal.f90:319-374).  The reference ships no vectors for these
routines, so they are pinned by what the mathematics fixes:

* each closure row is exact for polynomials up to the degree its free
  coefficients can match (3rd-order D1 row 1, 5th-order row 2; D2 rows
  exact to degree 4 and 6; the Visbal-Gaitonde filter rows to degree 3, the
  fourth row to 5), and fails one degree above -- a dropped term or a wrong
  coefficient breaks the lower degrees;
* the serial band solve agrees with a dense LU solve;
* the mirrored Gaussian equals numpy's reflect-padded convolution;
* the grid spans the period with both end points.
"""
import numpy as np
import pytest

import oracle as O
from fields import AXIS, along

NX = 16


@pytest.fixture(scope="module")
def onp():
    return O.Oracle(NX, 8, 8, period=(2.0, 2 * np.pi, 2 * np.pi), nonper_x=1)


def _a_dot(rows, w):
    """Rows (n, 5) = (lower-2 .. upper+2) of a band matrix times w."""
    n = len(w)
    out = np.zeros(n)
    for i in range(n):
        for k, off in enumerate((-2, -1, 0, 1, 2)):
            j = i + off
            if 0 <= j < n:
                out[i] += rows[i, k] * w[j]
    return out


def _dense(rows):
    n = len(rows)
    A = np.zeros((n, n))
    for i in range(n):
        for k, off in enumerate((-2, -1, 0, 1, 2)):
            if 0 <= i + off < n:
                A[i, i + off] = rows[i, k]
    return A


def _row_residual(o, op, row, p):
    """Residual of closure row `row` on the monomial ((x - x_row)/h)^p.

    The oracle returns A^-1 R(f); A times it recovers the row's right-hand
    side R(f)_row, which only reads points within 4 of the row, so f is the
    monomial there and zero elsewhere (keeping magnitudes small).  The exact
    left-hand side applies the row to the monomial's derivative."""
    n = o.params.nx
    h = o.params.period_x / (n - 1)
    xs = (np.arange(n) - row).astype(float)
    f = np.where(np.abs(xs) <= 4, xs ** p, 0.0)
    F = along(0, f, o.shape)
    rows = o.np_lhs(op)
    if op == 0:
        W = o.dfd(0, F)[0, 0, :] * h
        ex = p * xs ** (p - 1) if p > 0 else 0 * xs
    elif op == 1:
        W = o.d2fd2(0, F)[0, 0, :] * h * h
        ex = p * (p - 1) * xs ** (p - 2) if p > 1 else 0 * xs
    else:
        W = o.filter_dir(0, F)[0, 0, :]
        ex = xs ** p
    return (_a_dot(rows, W) - _a_dot(rows, ex))[row], 4.0 ** p


# exact degree per row from the near end: free coefficients - 1
EXACT = {
    0: [3, 5] + [6] * (NX - 4) + [5, 3],
    1: [4, 6] + [7] * (NX - 4) + [6, 4],
    2: [3, 3, 3, 5] + [7] * (NX - 8) + [5, 3, 3, 3],
}


@pytest.mark.parametrize("op", [0, 1, 2])
def test_closure_rows_exact_to_their_degree(onp, op):
    for row in range(NX):
        d = EXACT[op][row]
        for p in range(d + 1):
            r, scale = _row_residual(onp, op, row, p)
            assert abs(r) < 1e-8 * scale, (op, row, p, r)
        r, scale = _row_residual(onp, op, row, d + 1)
        assert abs(r) > 1e-3, (op, row, d + 1, r)


def test_lhs_rows(onp):
    """Boundary LHS rows as printed in Setup_Derivatives / Init_Filters."""
    d1, d2, fl = onp.np_lhs(0), onp.np_lhs(1), onp.np_lhs(2)
    np.testing.assert_array_equal(d1[0], [0, 0, 1, 2, 0])
    np.testing.assert_allclose(d1[1], [0, 1 / 6, 1, 1 / 2, 0])
    np.testing.assert_allclose(d1[NX // 2], [0, 1 / 3, 1, 1 / 3, 0])
    np.testing.assert_array_equal(d1[-1], d1[0][::-1])
    np.testing.assert_allclose(d2[1], [0, 0.1, 1, -7 / 20, 0])
    np.testing.assert_array_equal(d2[-2], d2[1][::-1])
    np.testing.assert_array_equal(fl[0], [0, 0, 1, 0, 0])
    np.testing.assert_allclose(fl[2], [0, 0.45, 1, 0.45, 0])
    np.testing.assert_allclose(fl[3], [0.16688, 0.66624, 1, 0.66624, 0.16688])


@pytest.mark.parametrize("n", [16, 33, 128])
def test_serial_band_solve_vs_dense(onp, n):
    rng = np.random.default_rng(n)
    # a diagonally dominant random band, and the operators' own rows
    rows = rng.uniform(-0.2, 0.2, size=(n, 5))
    rows[:, 2] = 1.0 + rng.uniform(0, 0.5, size=n)
    R = rng.normal(size=(3, n))
    X = O.np_band_solve(rows, R)
    for i in range(3):
        np.testing.assert_allclose(X[i], np.linalg.solve(_dense(rows), R[i]), rtol=0, atol=1e-13 * np.abs(X[i]).max())
    o = O.Oracle(n if n >= 16 else 16, 8, 8, period=(1.0, 1.0, 1.0), nonper_x=1) if n != NX else onp
    for op in range(3):
        rows = o.np_lhs(op)
        X = O.np_band_solve(rows, R)
        for i in range(3):
            ref = np.linalg.solve(_dense(rows), R[i])
            assert np.abs(X[i] - ref).max() <= 1e-11 * np.abs(ref).max()


def test_gaussian_mirror(onp):
    rng = np.random.default_rng(5)
    F = np.asfortranarray(rng.normal(size=onp.shape))
    a = np.array([107 / 103680, 149 / 12960, 1997 / 25920, 3091 / 12960, 3565 / 10368,
                  3091 / 12960, 1997 / 25920, 149 / 12960, 107 / 103680])
    Fp = np.pad(F, [(0, 0), (0, 0), (4, 4)], mode="reflect")   # f(-m) = f(m), f(n-1+m) = f(n-1-m)
    ref = sum(a[k] * Fp[:, :, k:k + NX] for k in range(9))
    np.testing.assert_allclose(onp.gauss_dir(0, F), ref, rtol=0, atol=1e-15 * np.abs(ref).max())


def test_grid_spans_period(onp):
    m = onp.metrics()
    assert m["D"][0] == pytest.approx(2.0 / (NX - 1), rel=1e-15)
    assert m["sx1"][0] == pytest.approx(1 / 2.0, rel=1e-15)
    assert m["D"][1] == pytest.approx(2 * np.pi / 8, rel=1e-15)


def test_periodic_directions_unchanged(onp):
    o = O.Oracle(NX, 8, 8, period=(2.0, 2 * np.pi, 2 * np.pi))
    F = np.asfortranarray(np.random.default_rng(2).normal(size=onp.shape))
    for d in (1, 2):
        np.testing.assert_array_equal(onp.dfd(d, F), o.dfd(d, F))
        np.testing.assert_array_equal(onp.filter_dir(d, F), o.filter_dir(d, F))
        np.testing.assert_array_equal(onp.gauss_dir(d, F), o.gauss_dir(d, F))


def test_d1_convergence():
    """D1 of sin on the closed interval converges in the max norm at the
    order of the third-order end row."""
    errs = []
    for n in (32, 64, 128):
        o = O.Oracle(n, 8, 8, period=(2.0, 1.0, 1.0), nonper_x=1)
        x = np.arange(n) * 2.0 / (n - 1)
        F = along(0, np.sin(2.1 * x + 0.3), o.shape)
        dF = o.dfd(0, F)[0, 0, :]
        errs.append(np.abs(dF - 2.1 * np.cos(2.1 * x + 0.3)).max())
    rates = np.log2(np.array(errs[:-1]) / np.array(errs[1:]))
    assert ((rates > 2.8) & (rates < 3.3)).all(), rates


def test_static_hyper_and_first_step_limit():
    """src/hyper_routines.f:99-101,204-206: beta* = mu*/C_mu; the dt limit
    (src/lddrk4_nonperiodic_sf.f:105) applies from the first step."""
    from fields import smooth_state
    shape = (32, 16, 16)
    o = O.Oracle(*shape, period=(2 * np.pi,) * 3, Re=80.0, nonper_x=1)
    Q = smooth_state(*shape, seed=3)
    _, cfl = o.cfl(Q)
    assert o.tstep(Q, cfl, 1) == 0.5 / cfl          # coefficients still zero: limits infinite
    o.rhs(Q)
    o.compute_hyper(Q)
    H = o.hyper()
    np.testing.assert_allclose(H[..., 0], H[..., 1] / 2.0e-3, rtol=1e-15)
    assert (H[..., 2] > 0).all()
    Hb = np.asfortranarray(H * 1e6)                   # large enough to bind
    o.set_hyper(Hb)
    assert o.tstep(Q, cfl, 1) < 0.5 / cfl
    p = O.Oracle(*shape, period=(2 * np.pi,) * 3, Re=80.0)
    p.set_hyper(Hb)
    assert p.tstep(Q, cfl, 1) == 0.5 / cfl           # the periodic driver waits for step 2
    assert p.tstep(Q, cfl, 2) < 0.5 / cfl
