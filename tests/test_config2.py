"""Config-2 pieces on the host and in the oracle (CPU): the planar shock
initial condition obeys the Rankine-Hugoniot jump conditions, and the
oracle's Sponge_Forcing (src/sponge1d.f) has the reference's strength
profile."""
import numpy as np

import oracle as O
import st_amd
from fields import smooth_state


def test_shock_ic_rankine_hugoniot():
    g, M = 1.4, 1.5
    nx, ny, nz = 64, 16, 16
    Q, s1, s2 = st_amd.ic_shock(nx, ny, nz, (4.0, 1.0, 1.0), mach=M, frac=0.25, urms=0.0, gamma=g)

    def fluxes(s):
        rho, u = s[3], s[0] / s[3]
        p = (g - 1) * (s[4] - 0.5 * rho * u * u)
        return np.array([rho * u, rho * u * u + p, (s[4] + p) * u]), rho, u, p

    f1, r1, u1, p1 = fluxes(s1)
    f2, r2, u2, p2 = fluxes(s2)
    np.testing.assert_allclose(f2, f1, rtol=1e-14)             # mass, momentum, energy fluxes
    assert abs(u1 / np.sqrt(g * p1 / r1) - M) < 1e-14         # upstream Mach number
    assert r2 > r1 and p2 > p1 and u2 < u1                      # compressive
    # far upstream and downstream (the tanh tail is 1e-7 at 8 widths)
    np.testing.assert_allclose(Q[0, 0, 0], s1, rtol=1e-6, atol=1e-12)
    np.testing.assert_allclose(Q[0, 0, -1], s2, rtol=1e-6, atol=1e-12)


def test_sponge_strength_profile():
    nx, ny, nz = 40, 8, 8
    per = (2.0, 1.0, 1.0)
    Q = smooth_state(nx, ny, nz, seed=3)
    refl = np.array([0.6, 0.0, 0.0, 1.05, 2.4])
    refr = np.array([0.3, 0.0, 0.0, 1.80, 3.1])
    out = []
    for amp in (0.0, 5.0):
        o = O.Oracle(nx, ny, nz, period=per, nonper_x=1)
        o.set_sponge(amp, 2.0, 0.4, refl, refr)
        out.append(o.rk_stage(Q, np.zeros_like(Q), 0.0, 0.1, 1e-3)[2])
    d = out[1] - out[0]                                        # the forcing alone
    ref = np.where(np.arange(nx)[None, None, :, None] < 20, refl, refr)
    sig = -d / (Q - ref)                                       # sigma per point
    s = sig[2, 3, :, 1]
    np.testing.assert_allclose(sig, np.broadcast_to(s[None, None, :, None], sig.shape), rtol=1e-6, atol=1e-9)
    left = int(np.ceil(0.4 * nx / 2))                          # 8 points per layer
    assert np.all(s[left:nx - left] == 0) and s[left - 1] > 0 and s[nx - left] > 0 and s[nx - left - 1] == 0
    k = np.arange(1, left + 1)
    np.testing.assert_allclose(s[:left], 5.0 * ((left - k + 1) / left) ** 2, rtol=1e-6)
    np.testing.assert_allclose(s[::-1][:left - 1], s[:left - 1], rtol=1e-6)   # symmetric layers
