"""Pins the oracle's composite physics without the reference.

RHS (src/rhsv_periodic_sf.f) and Hyper_RHS (src/hyper_rhs.f):
  * a uniform state has zero right-hand side;
  * the continuity flux is a compact first derivative, so the domain sum
    of d(rho)/dt vanishes to roundoff; every hyper term is a divergence,
    so the hyper momentum and energy sums vanish too;
  * swapping x and y (and the two momentum components) swaps the result;
  * against the continuous Navier-Stokes right-hand side evaluated with
    spectral derivatives, the error falls at 6th order under refinement
    (a dropped term or a wrong sign or index shows up as an O(1) error).
Time integration (src/lddrk4_periodic_sf.f):
  * RK tables are 4th order on a linear ODE;
  * one stage equals RHS, Hyper_RHS, update, filter, pressure fix in turn;
  * Fix_Pressure clamps negative pressure to 0.71463;
  * Compute_CFL / Compute_Tstep closed forms on a uniform state.
"""
import numpy as np
import pytest

import oracle as O
from fields import coords, relerr, smooth_state, spectral_deriv

GAMMA = 1.4


def test_uniform_state_zero():
    n = 16
    o = O.Oracle(n, n, n, Re=100.0)
    Q = np.zeros((n, n, n, 5), order="F")
    Q[..., 3] = 1.3
    Q[..., 0] = 0.4
    Q[..., 1] = -0.2
    Q[..., 4] = 7.0
    r = o.full_rhs(Q)
    assert np.abs(r).max() < 1e-12


@pytest.mark.parametrize("shape", [(16, 16, 16), (24, 16, 32)])
def test_conservation(shape):
    nx, ny, nz = shape
    o = O.Oracle(nx, ny, nz, Re=50.0)
    Q = smooth_state(nx, ny, nz, seed=3)
    r, _ = o.rhs(Q)
    assert abs(r[..., 3].sum()) < 1e-12 * np.abs(r[..., 3]).sum()
    h = o.hyper_rhs(Q, np.zeros_like(r))
    for c in (0, 1, 2, 4):
        assert abs(h[..., c].sum()) < 1e-11 * max(np.abs(h[..., c]).sum(), 1e-300)
    assert np.abs(h[..., 3]).max() == 0.0


def test_axis_swap_symmetry():
    n = 16
    o = O.Oracle(n, n, 12, Re=40.0)
    Q = smooth_state(n, n, 12, seed=5)
    # swap x <-> y: array axes 0 (y) and 2 (x), momentum components 0 and 1
    Qs = np.asfortranarray(np.transpose(Q, (2, 1, 0, 3))[..., [1, 0, 2, 3, 4]])
    r = o.full_rhs(Q)
    rs = o.full_rhs(Qs)
    back = np.asfortranarray(np.transpose(rs, (2, 1, 0, 3))[..., [1, 0, 2, 3, 4]])
    assert relerr(back, r).max() < 1e-12


def _continuous_rhs(Q, Re, Pr, vis_exp, bulk, gamma=GAMMA):
    """Navier-Stokes RHS in conservative form with spectral derivatives."""
    ax = {0: 2, 1: 0, 2: 1}
    D = lambda f, d: spectral_deriv(f, ax[d])
    rho = Q[..., 3]
    u = [Q[..., i] / rho for i in range(3)]
    p = (gamma - 1) * (Q[..., 4] - 0.5 * (Q[..., 0] ** 2 + Q[..., 1] ** 2 + Q[..., 2] ** 2) / rho)
    T = gamma / (gamma - 1) * p / rho
    mu = ((gamma - 1) * T) ** vis_exp / Re
    gu = [[D(u[i], j) for j in range(3)] for i in range(3)]
    div = gu[0][0] + gu[1][1] + gu[2][2]
    tau = [[mu * (gu[i][j] + gu[j][i]) + ((bulk - 2.0 / 3.0) * mu * div if i == j else 0.0)
            for j in range(3)] for i in range(3)]
    out = np.zeros_like(Q)
    out[..., 3] = -sum(D(rho * u[j], j) for j in range(3))
    for i in range(3):
        out[..., i] = sum(-D(rho * u[i] * u[j], j) + D(tau[i][j], j) for j in range(3)) - D(p, i)
    E = Q[..., 4]
    out[..., 4] = sum(-D((E + p) * u[j], j)
                      + D(sum(u[i] * tau[i][j] for i in range(3)), j)
                      + D(mu / Pr * D(T, j), j) for j in range(3))
    return out


def _mms_state(n):
    X, Y, Z = coords(n, n, n)
    rho = 1.0 + 0.1 * np.sin(X + Y) * np.cos(Z)
    u = 0.5 * np.sin(X) * np.cos(Y)
    v = -0.5 * np.cos(X) * np.sin(Y) + 0.2 * np.sin(Z)
    w = 0.3 * np.cos(X + Z)
    p = 5.0 + 0.3 * np.cos(Y - Z)
    Q = np.zeros((n, n, n, 5), order="F")
    Q[..., 0], Q[..., 1], Q[..., 2], Q[..., 3] = rho * u, rho * v, rho * w, rho
    Q[..., 4] = p / (GAMMA - 1) + 0.5 * rho * (u * u + v * v + w * w)
    return Q


@pytest.mark.parametrize("viscous", [False, True])
def test_mms_sixth_order(viscous):
    Re, Pr = (20.0, 0.7) if viscous else (0.0, 0.0)
    errs = []
    for n in (16, 32):
        o = O.Oracle(n, n, n, Re=Re, Pr=Pr, vis_exp=0.76, bulk_ratio=0.3, hyper_param=0)
        Q = _mms_state(n)
        r, _ = o.rhs(Q)
        ex = _continuous_rhs(Q, Re if viscous else np.inf, Pr if viscous else 1.0, 0.76, 0.3)
        errs.append(np.abs(r - ex).reshape(-1, 5).max(axis=0))
    errs = np.array(errs)
    ratio = errs[0] / errs[1]
    assert ratio.min() > 40.0, (errs, ratio)     # 2^6 = 64


def test_rk_tables_fourth_order():
    for scheme in (0, 1):
        errs = []
        for nsteps in (10, 20):
            lam = -1.0 + 2.0j
            dt = 1.0 / nsteps / 2.0          # one step = two RK sweeps of dt
            y = 1.0 + 0j
            for _ in range(nsteps):
                for loop in (0, 1):
                    a, b = O.rk_coeffs(scheme, loop)
                    qw = 0.0
                    for k in range(len(a)):
                        qw = a[k] * qw + dt * lam * y
                        y = y + b[k] * qw
            errs.append(abs(y - np.exp(lam)))
        assert 14.0 < errs[0] / errs[1] < 18.5, errs


def test_stage_composition():
    n = 16
    o = O.Oracle(n, n, n, Re=60.0)
    Q = smooth_state(n, n, n, seed=7)
    QW = 0.01 * smooth_state(n, n, n, seed=8)
    a, b, dt = -0.48, 0.74, 2e-3
    Qn, QWn, dQ = o.rk_stage(Q, QW, a, b, dt)
    r = o.full_rhs(Q)
    np.testing.assert_array_equal(dQ, r)
    QW2 = a * QW + dt * r
    np.testing.assert_array_equal(QWn, QW2)
    Q2 = o.fix_pressure(o.compact_filter(np.asfortranarray(Q + b * QW2)))
    np.testing.assert_array_equal(Qn, Q2)


def test_fix_pressure_clamp():
    n = 8
    o = O.Oracle(n, n, n)
    Q = smooth_state(n, n, n, seed=1)
    Q[0, 0, 0, 4] = 0.0                       # negative pressure here
    out = o.fix_pressure(Q)
    ke = 0.5 * (Q[..., 0] ** 2 + Q[..., 1] ** 2 + Q[..., 2] ** 2) / Q[..., 3]
    assert abs(out[0, 0, 0, 4] - (0.71463 / 0.4 + ke[0, 0, 0])) < 1e-14
    p = 0.4 * (Q[..., 4] - ke)
    mask = np.ones_like(p, dtype=bool)
    mask[0, 0, 0] = False
    np.testing.assert_allclose(out[..., 4][mask], (p / 0.4 + ke)[mask], rtol=1e-14)
    np.testing.assert_array_equal(out[..., :4], Q[..., :4])


def test_cfl_and_tstep_uniform():
    nx, ny, nz = 16, 12, 8
    o = O.Oracle(nx, ny, nz, cfl_cal=0.5)
    Q = np.zeros((ny, nz, nx, 5), order="F")
    rho, u0, p = 1.2, 0.7, 3.0
    Q[..., 3] = rho
    Q[..., 0] = rho * u0
    Q[..., 4] = p / 0.4 + 0.5 * rho * u0 * u0
    stab, resv = o.cfl(Q)
    c = np.sqrt(GAMMA * p / rho)
    D = o.metrics()["D"]
    exp_resv = (c + u0) * 2.199 / D[0] + c * 2.199 / D[1] + c * 2.199 / D[2]
    exp_stab = (c + u0) * 2.478 / D[0] + c * 2.478 / D[1] + c * 2.478 / D[2]
    assert abs(resv - exp_resv) < 1e-12 * exp_resv
    assert abs(stab - exp_stab) < 1e-12 * exp_stab
    assert o.tstep(Q, resv, 0) == 0.5 / resv
    # hyper limits only from step 2 on; a uniform state has zero coefficients (Inf limits)
    o.compute_hyper(Q)
    assert o.tstep(Q, resv, 5) == 0.5 / resv
    H = np.full((ny, nz, nx, 3), 1e-3, order="F")
    o.set_hyper(H)
    dt = o.tstep(Q, resv, 5)
    T = GAMMA * (p / 0.4) / rho
    lim = min(rho * D[0] ** 2 / 1e-3, rho * (D[0] * c) ** 2 / (1e-3 * T))
    assert abs(dt - min(0.5 / resv, float(np.float32(0.2)) * lim)) < 1e-15


def test_diagnostics_uniform():
    n = 8
    o = O.Oracle(n, n, n)
    Q = np.zeros((n, n, n, 5), order="F")
    Q[..., 3] = 1.1
    Q[..., 1] = 0.33
    Q[..., 4] = 6.0
    d = o.diagnostics(Q)
    p = 0.4 * (6.0 - 0.5 * 0.33 ** 2 / 1.1)
    np.testing.assert_allclose(d, [0.0, 0.33, 0.0, 0.5 * 0.33 ** 2 / 1.1, 1.1, 6.0,
                                   1.1 * np.log(p * (1 / 1.1) ** GAMMA)], rtol=1e-13, atol=1e-15)
