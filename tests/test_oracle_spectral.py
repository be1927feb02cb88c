"""Pins the oracle's composites to rounding with a second, independent
restatement (tests/spectral_ref.py: the compact operators applied through
their exact discrete symbols in Fourier space), plus closed forms of the
dilatation sensor.

Covered oracle functions: RHS and /GRAD/ (src/rhsv_periodic_sf.f),
Compute_coeff1 / Compute_hyper (src/hyper_dynamic.f:72-318), Hyper_RHS
(src/hyper_rhs.f:12-499), one LDDRK4 stage with the filter and
Fix_Pressure (src/lddrk4_periodic_sf.f:344-356, :436-473).

A plausible slip in the oracle (mu* and beta* swapped, Delta^5 and
Delta^6 swapped, a wrong C_k or C_mu, a wrong eta_2 factor, a sign or a
factor of the tanh sensor, a dropped viscous cross term) moves these
results by O(1) relative, far above the 1e-13 bars (measured agreement: 1e-15 to 8e-15); the planted-mistake
test below shows the comparison catches such a slip.
"""
import numpy as np
import pytest

import oracle as O
from fields import relerr, smooth_state, sym_d1
from spectral_ref import Spectral

SHAPES = [(16, 16, 16), (32, 24, 16), (24, 16, 32)]


def _pair(shape, **kw):
    nx, ny, nz = shape
    return O.Oracle(nx, ny, nz, **kw), Spectral(nx, ny, nz, **{k: v for k, v in kw.items() if k != "hyper_param"})


def _state(shape, seed, amp=1.0, mt=0.3):
    return smooth_state(*shape, seed=seed, amp=amp, mt=mt, nmodes=8)


@pytest.mark.parametrize("bulk", [0.0, 0.7])
@pytest.mark.parametrize("shape", SHAPES)
def test_rhs_matches_spectral(shape, bulk):
    o, s = _pair(shape, Re=60.0, Pr=0.7, bulk_ratio=bulk, hyper_param=0)
    Q = _state(shape, seed=5)
    r_o, _ = o.rhs(Q)
    r_s, G_s = s.rhs(Q)
    assert relerr(r_o, r_s).max() < 1e-13
    assert relerr(o.grad(), G_s).max() < 1e-13


@pytest.mark.parametrize("shape", SHAPES)
def test_hyper_composites_match_spectral(shape):
    o, s = _pair(shape, Re=60.0, Pr=0.7, hyper_param=1)
    Q = _state(shape, seed=9, amp=1.5, mt=0.5)
    r_o, _ = o.rhs(Q)
    full_o = o.hyper_rhs(Q, r_o)
    H_o = o.hyper()
    r_s, G_s = s.rhs(Q)
    H_s = s.hyper(Q, G_s)
    eh = relerr(H_o, H_s)
    assert eh.max() < 1e-13, eh
    # the artificial-transport terms alone, then the whole right-hand side
    Qh_s = s.hyper_rhs(Q, G_s, H_s)
    eq = relerr(full_o - r_o, Qh_s)
    assert eq.max() < 1e-13, eq
    assert relerr(full_o, r_s + Qh_s).max() < 1e-13


@pytest.mark.parametrize("shape", SHAPES[:2])
def test_stage_matches_spectral(shape):
    o, s = _pair(shape, Re=60.0, Pr=0.7, hyper_param=1)
    Q = _state(shape, seed=3, amp=1.0, mt=0.4)
    QW = np.asfortranarray(0.01 * _state(shape, seed=4))
    a, b, dt = -0.4801594388477952, 0.7408540575766301, 2e-3
    Qo, QWo, _ = o.rk_stage(Q, QW, a, b, dt)
    Qs, QWs = s.stage(Q, QW, a, b, dt)
    assert relerr(QWo, QWs).max() < 1e-13
    # the filter's band solve (condition ~2e3) sets the stage's spread
    assert relerr(Qo, Qs).max() < 1e-12


def _sensor(dil):
    """C_beta of src/hyper_dynamic.f:142-163 with omega = 0."""
    fac = 0.5 * (1.0 - np.tanh(2.5 + 10.0 * dil))
    return fac * dil * dil / (dil * dil + 1.0e-7) + 1.0e-20


def _cbeta_from_oracle(o, Q):
    """beta* / mu* = C_beta / C_mu exactly when eta_3 = eta_1 (src/hyper_dynamic.f:286-314)."""
    r, _ = o.rhs(Q)
    o.hyper_rhs(Q, r)
    H = o.hyper()
    return 2.0e-3 * H[..., 0] / H[..., 1]


@pytest.mark.parametrize("amp", [0.05, 0.3])
def test_sensor_closed_form_irrotational(amp):
    """u = amp cos(k x), v = w = 0, rho = 1: omega vanishes, dil = -amp k~ sin(k x)."""
    n = 32
    o = O.Oracle(n, 16, 16, Re=100.0, Pr=0.7, hyper_param=1)
    h = 2 * np.pi / n
    x = np.arange(n) * h
    k = 3
    shape = (16, 16, n)
    u = np.broadcast_to(amp * np.cos(k * x)[None, None, :], shape)
    Q = np.zeros(shape + (5,), order="F")
    Q[..., 0] = u
    Q[..., 3] = 1.0
    Q[..., 4] = 1.0 / (1.4 * 0.09) / 0.4 + 0.5 * u * u
    dil = -amp * sym_d1(k * h) / h * np.sin(k * x)
    want = np.broadcast_to(_sensor(dil)[None, None, :], shape)
    got = _cbeta_from_oracle(o, Q)
    assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()
    # and the sign convention of the switch: compression (dil < 0) turns the
    # sensor on, expansion (dil > 0) turns it off
    i_c = np.argmin(dil)
    i_e = np.argmax(dil)
    assert got[0, 0, i_c] > 100.0 * got[0, 0, i_e]


def test_sensor_closed_form_solenoidal():
    """u = sin(y), v = sin(z), w = sin(x): each velocity is constant along its
    own direction, so the discrete dilatation vanishes and C_beta = 1e-20."""
    n = 16
    o = O.Oracle(n, n, n, Re=100.0, Pr=0.7, hyper_param=1)
    x = np.arange(n) * 2 * np.pi / n
    X, Y, Z = x[None, None, :], x[:, None, None], x[None, :, None]
    shape = (n, n, n)
    Q = np.zeros(shape + (5,), order="F")
    Q[..., 0] = np.broadcast_to(np.sin(Y), shape)
    Q[..., 1] = np.broadcast_to(np.sin(Z), shape)
    Q[..., 2] = np.broadcast_to(np.sin(X), shape)
    Q[..., 3] = 1.0
    Q[..., 4] = 10.0 + 0.5 * (Q[..., 0] ** 2 + Q[..., 1] ** 2 + Q[..., 2] ** 2)
    cb = _cbeta_from_oracle(o, Q)
    np.testing.assert_allclose(cb, 1.0e-20, rtol=1e-12)


def test_planted_mistake_is_caught():
    """Swapping mu* and beta* in a copy of the oracle's result (the check
    sees a mistake in the oracle) fails the composite comparison."""
    shape = (16, 16, 16)
    o, s = _pair(shape, Re=60.0, Pr=0.7, hyper_param=1)
    Q = _state(shape, seed=9, amp=1.5, mt=0.5)
    r_o, _ = o.rhs(Q)
    o.hyper_rhs(Q, r_o)
    H_o = o.hyper()
    _, G_s = s.rhs(Q)
    H_s = s.hyper(Q, G_s)
    bad = H_o[..., [1, 0, 2]]
    assert relerr(bad, H_s).max() > 1e-3
    bad = H_o.copy(order="F")
    bad[..., 2] *= (2 * np.pi / 16)        # Delta^6 in place of Delta^5
    assert relerr(bad, H_s).max() > 1e-3
