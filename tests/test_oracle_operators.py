"""Pins the oracle's compact operators and band solver without the reference.

* sinusoids are eigenvectors of the periodic (circulant) operators, so D1,
  D2 (both coefficient mappings), the compact filter and the Gaussian are
  pinned to roundoff by their closed-form symbols;
* refinement shows 6th-order convergence of D1 and D2;
* the cyclic band solver (src/pentadiagonal.f90:381-504) is checked
  against a dense solve on tiny lines;
* the one known-answer check the reference itself carries,
  DFDX(sin x) - cos x (src/par_test_stuff.f:16-26), is repeated.
"""
import numpy as np
import pytest

import oracle as O
from fields import (AXIS, along, grid_1d, sym_d1, sym_d2, sym_filter, sym_gauss,
                    FILT_AL, FILT_BE)

SHAPES = [(16, 16, 16), (24, 16, 32), (32, 24, 16)]


@pytest.mark.parametrize("nx,ny,nz", SHAPES)
@pytest.mark.parametrize("d", [0, 1, 2])
def test_d1_symbol(nx, ny, nz, d):
    o = O.Oracle(nx, ny, nz)
    n = (nx, ny, nz)[d]
    h = 2 * np.pi / n
    x = grid_1d(n)
    for k in (1, 2, n // 4, n // 2 - 1):
        F = along(d, np.sin(k * x), (ny, nz, nx))
        dF = o.dfd(d, F)
        exact = along(d, sym_d1(k * h) / h * np.cos(k * x), (ny, nz, nx))
        assert np.abs(dF - exact).max() < 1e-13 * max(1.0, k)
        # accumulate with a scale (iadd=1, cadd) as the RHS uses it
        acc = o.dfd(d, F, dF=np.ones((ny, nz, nx)), iadd=1, cadd=-0.5)
        assert np.abs(acc - (1.0 - 0.5 * exact)).max() < 1e-13 * max(1.0, k)


@pytest.mark.parametrize("nx,ny,nz", SHAPES)
@pytest.mark.parametrize("d", [0, 1, 2])
@pytest.mark.parametrize("shipped", [False, True])
def test_d2_symbol(nx, ny, nz, d, shipped):
    o = O.Oracle(nx, ny, nz, d2_mode=1 if shipped else 0)
    n = (nx, ny, nz)[d]
    h = 2 * np.pi / n
    x = grid_1d(n)
    for k in (0, 1, 3, n // 2 - 2):
        F = along(d, np.cos(k * x), (ny, nz, nx))
        dF, d2F = o.d2fd2(d, F, iadd=1)
        exact = along(d, -sym_d2(k * h, shipped) / h ** 2 * np.cos(k * x), (ny, nz, nx))
        scale = max(1.0, np.abs(exact).max())
        assert np.abs(d2F - exact).max() < 2e-13 * scale
        ex1 = along(d, -sym_d1(k * h) / h * np.sin(k * x), (ny, nz, nx))
        assert np.abs(dF - ex1).max() < 1e-13 * max(1.0, k)


def test_d2_constant():
    """Consistent mapping annihilates constants; the shipped one gives 5.1 f/h^2."""
    n = 16
    h = 2 * np.pi / n
    F = np.full((n, n, n), 2.0, order="F")
    o = O.Oracle(n, n, n, d2_mode=0)
    assert np.abs(o.d2fd2(0, F)).max() < 1e-12
    o1 = O.Oracle(n, n, n, d2_mode=1)
    np.testing.assert_allclose(o1.d2fd2(1, F), 5.1 * 2.0 / h ** 2, rtol=1e-13)


@pytest.mark.parametrize("d", [0, 1, 2])
def test_sixth_order(d):
    errs1, errs2 = [], []
    for n in (16, 32, 64):
        o = O.Oracle(n if d == 0 else 8, n if d == 1 else 8, n if d == 2 else 8)
        shape = o.shape
        x = grid_1d(n)
        f = np.exp(np.sin(x))
        F = along(d, f, shape)
        dF, d2F = o.d2fd2(d, F, iadd=1)
        ex1 = np.cos(x) * f
        ex2 = (np.cos(x) ** 2 - np.sin(x)) * f
        errs1.append(np.abs(dF - along(d, ex1, shape)).max())
        errs2.append(np.abs(d2F - along(d, ex2, shape)).max())
    r1 = np.log2(errs1[0] / errs1[1]), np.log2(errs1[1] / errs1[2])
    r2 = np.log2(errs2[0] / errs2[1]), np.log2(errs2[1] / errs2[2])
    assert min(r1) > 5.7, r1
    assert min(r2) > 5.7, r2


@pytest.mark.parametrize("d", [0, 1, 2])
def test_filter_symbol(d):
    nx, ny, nz = 32, 16, 24
    o = O.Oracle(nx, ny, nz)
    n = (nx, ny, nz)[d]
    h = 2 * np.pi / n
    x = grid_1d(n)
    for k in (0, 1, 5, n // 2 - 1, n // 2):
        F = along(d, np.cos(k * x), (ny, nz, nx))
        out = o.filter_dir(d, F)
        exact = along(d, sym_filter(k * h) * np.cos(k * x), (ny, nz, nx))
        assert np.abs(out - exact).max() < 2e-13
        g = o.gauss_dir(d, F)
        exact_g = along(d, sym_gauss(k * h) * np.cos(k * x), (ny, nz, nx))
        assert np.abs(g - exact_g).max() < 1e-15 * 8
    # constants preserved, Nyquist annihilated (closed form: T(0)=1, T(pi)=0)
    F = np.full((ny, nz, nx), 3.0, order="F")
    assert np.abs(o.filter_dir(d, F) - 3.0).max() < 3.0 * 2e-13   # cond(A) ~ 2e3
    assert np.abs(o.gauss_dir(d, F) - 3.0).max() < 1e-14
    nyq = along(d, np.cos(np.pi * np.arange(n)), (ny, nz, nx))
    assert np.abs(o.filter_dir(d, nyq)).max() < 1e-11
    assert np.abs(o.gauss_dir(d, nyq)).max() < 1e-15


def test_filter_3d_product_symbol():
    """Compact_Filter applies X, Y, Z in sequence: symbol T(tx)T(ty)T(tz)."""
    nx, ny, nz = 32, 16, 24
    o = O.Oracle(nx, ny, nz)
    from fields import coords
    X, Y, Z = coords(nx, ny, nz)
    kx, ky, kz = 3, 2, 5
    F = np.asfortranarray(np.cos(kx * X) * np.cos(ky * Y) * np.cos(kz * Z))
    Q = np.asfortranarray(np.stack([F, 2 * F], axis=3))
    out = o.compact_filter(Q)
    s = sym_filter(kx * 2 * np.pi / nx) * sym_filter(ky * 2 * np.pi / ny) * sym_filter(kz * 2 * np.pi / nz)
    assert np.abs(out[..., 0] - s * F).max() < 3e-13          # three passes, cond ~2e3 each
    assert np.abs(out[..., 1] - 2 * s * F).max() < 6e-13
    g = o.gaussian_filter(Q)
    sg = sym_gauss(kx * 2 * np.pi / nx) * sym_gauss(ky * 2 * np.pi / ny) * sym_gauss(kz * 2 * np.pi / nz)
    assert np.abs(g[..., 0] - sg * F).max() < 1e-14


@pytest.mark.parametrize("n", list(range(5, 17)))
@pytest.mark.parametrize("ab", [(1.0 / 3.0, 0.0), (2.0 / 11.0, 0.0), (FILT_AL, FILT_BE)])
def test_band_solver_dense(n, ab):
    al, be = ab
    rng = np.random.default_rng(n)
    R = rng.normal(size=(4, n))
    X = O.band_solve(R, al, be)
    A = np.zeros((n, n))
    for i in range(n):
        A[i, i] += 1.0
        A[i, (i + 1) % n] += al
        A[i, (i - 1) % n] += al
        A[i, (i + 2) % n] += be
        A[i, (i - 2) % n] += be
    Xd = np.linalg.solve(A, R.T).T
    cond = np.linalg.cond(A)
    assert np.abs(X - Xd).max() <= 20 * cond * 1e-16 * np.abs(Xd).max()


def test_known_answer_sin():
    """src/par_test_stuff.f:16-26 prints DFDX(sin x) - cos x per rank."""
    n = 64
    o = O.Oracle(n, 8, 8)
    x = grid_1d(n)
    F = along(0, np.sin(x), o.shape)
    err = np.abs(o.dfd(0, F) - along(0, np.cos(x), o.shape)).max()
    h = 2 * np.pi / n
    assert abs(err - abs(sym_d1(h) / h - 1.0)) < 1e-13
    assert err < 1e-9
