"""Invariants of the seeded isotropic-turbulence initial condition
(st_ic_isotropic, host side; no GPU).

The workloads are defined by it: a random-phase solenoidal velocity with
shell spectrum E(k) ~ k^4 exp(-2 (k/k0)^2), per-component rms
sqrt(<u.u>/3) = u_rms, rho = 1 and p = 1/(gamma M_t^2), bit-identical per
seed.  Checked with numpy FFTs, independently of the generator's own
transform.
"""
import hashlib
import os
import subprocess
import sys

import numpy as np
import pytest

import st_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _wavenumbers(shape, period):
    ny, nz, nx = shape
    ky = 2 * np.pi / period[1] * np.fft.fftfreq(ny, 1.0 / ny)
    kz = 2 * np.pi / period[2] * np.fft.fftfreq(nz, 1.0 / nz)
    kx = 2 * np.pi / period[0] * np.fft.fftfreq(nx, 1.0 / nx)
    return ky[:, None, None], kz[None, :, None], kx[None, None, :]


CASES = [((64, 64, 64), 4.0, 0.3, 12345), ((64, 64, 64), 8.0, 0.5, 7), ((48, 32, 64), 3.0, 0.2, 21)]


@pytest.mark.parametrize("shape_xyz,k0,mt,seed", CASES)
def test_ic_solenoidal_rms_and_thermo(shape_xyz, k0, mt, seed):
    nx, ny, nz = shape_xyz
    g = 1.4
    Q = st_amd.ic_isotropic(nx, ny, nz, k0=k0, urms=1.0, mt=mt, seed=seed)
    u, v, w = Q[..., 0], Q[..., 1], Q[..., 2]
    # rho = 1 exactly; p = 1/(gamma Mt^2) to rounding
    assert np.all(Q[..., 3] == 1.0)
    p = (g - 1.0) * (Q[..., 4] - 0.5 * (u * u + v * v + w * w) / Q[..., 3])
    np.testing.assert_allclose(p, 1.0 / (g * mt * mt), rtol=1e-13)
    # per-component rms convention: sqrt(<u.u>/3) = u_rms
    assert abs(np.sqrt(np.mean(u * u + v * v + w * w) / 3.0) - 1.0) < 1e-12     # summation rounding over 1e5 points
    # spectral divergence: max |k . u_hat| at rounding relative to max |k| |u_hat|
    ky, kz, kx = _wavenumbers(u.shape, (2 * np.pi,) * 3)
    U, V, W = (np.fft.fftn(f) for f in (u, v, w))
    div = np.abs(kx * U + ky * V + kz * W).max()
    scale = np.sqrt(kx ** 2 + ky ** 2 + kz ** 2).max() * max(np.abs(U).max(), np.abs(V).max(), np.abs(W).max())
    assert div < 1e-12 * scale
    # no mode at or beyond the cut-off radius N/2 (grid units)
    kk = np.sqrt(kx ** 2 + ky ** 2 + kz ** 2)
    e = np.abs(U) ** 2 + np.abs(V) ** 2 + np.abs(W) ** 2
    assert e[kk >= 0.5 * min(nx, ny, nz)].max() < 1e-24 * e.max()
    assert e[0, 0, 0] < 1e-24 * e.max()                         # zero mean velocity


@pytest.mark.parametrize("k0,seed", [(4.0, 12345), (8.0, 7), (4.0, 99)])
def test_ic_shell_spectrum(k0, seed):
    """Shell energies follow k^4 exp(-2 (k/k0)^2) within sampling noise.

    Each shell's energy sums ~N_s/2 independent Hermitian pairs of
    2-component complex Gaussian amplitudes, so its relative spread is
    about 1/sqrt(N_s); every shell with enough modes lies within 5 of
    those, and a spectrum of another shape (k^2 exp(-(k/k0)^2), say) does
    not.
    """
    n = 64
    Q = st_amd.ic_isotropic(n, n, n, k0=k0, urms=1.0, mt=0.3, seed=seed)
    ky, kz, kx = _wavenumbers((n, n, n), (2 * np.pi,) * 3)
    kk = np.sqrt(kx ** 2 + ky ** 2 + kz ** 2)
    e = sum(np.abs(np.fft.fftn(Q[..., c])) ** 2 for c in range(3))
    shell = np.rint(kk).astype(int)
    ok = (kk > 0) & (kk < n / 2)

    def ratios(model):
        # expected shell energy: per-mode density E(k)/(4 pi k^2) summed over
        # the shell's modes, with one overall constant fitted to the total
        per_mode = model(kk) / (kk ** 2 + (kk == 0))
        s_obs = np.bincount(shell[ok], weights=e[ok])
        s_exp = np.bincount(shell[ok], weights=per_mode[ok])
        cnt = np.bincount(shell[ok])
        use = (cnt >= 30) & (s_exp > 1e-6 * s_exp.max())
        c = s_obs[use].sum() / s_exp[use].sum()
        return s_obs[use] / (c * s_exp[use]), cnt[use]

    r, cnt = ratios(lambda k: k ** 4 * np.exp(-2 * (k / k0) ** 2))
    assert len(r) >= 8
    assert np.all(np.abs(r - 1.0) < 5.0 / np.sqrt(cnt)), np.c_[r, cnt]
    # a spectrum of another shape is rejected by the same test
    r_bad, cnt_bad = ratios(lambda k: k ** 2 * np.exp(-(k / k0) ** 2))
    assert not np.all(np.abs(r_bad - 1.0) < 5.0 / np.sqrt(cnt_bad))


def test_ic_seed_changes_field():
    a = st_amd.ic_isotropic(32, 32, 32, k0=4.0, urms=1.0, mt=0.3, seed=1)
    b = st_amd.ic_isotropic(32, 32, 32, k0=4.0, urms=1.0, mt=0.3, seed=2)
    c = st_amd.ic_isotropic(32, 32, 32, k0=4.0, urms=1.0, mt=0.3, seed=1)
    assert np.array_equal(a, c)
    assert np.abs(a - b).max() > 0.1


_HASH = """
import hashlib, sys
sys.path.insert(0, {py!r})
import st_amd
Q = st_amd.ic_isotropic(48, 40, 32, k0=6.0, urms=1.0, mt=0.5, seed=4242)
print(hashlib.sha256(Q.tobytes(order="F")).hexdigest())
"""


def test_ic_bit_identical_across_threads():
    py = os.path.join(ROOT, "shock-turbulence_amd", "python")
    out = []
    for t in ("1", "3", "8"):
        env = dict(os.environ, OMP_NUM_THREADS=t)
        r = subprocess.run([sys.executable, "-c", _HASH.format(py=py)], env=env, capture_output=True, text=True,
                           timeout=300)
        assert r.returncode == 0, r.stderr
        out.append(r.stdout.strip())
    assert out[0] == out[1] == out[2]
    Q = st_amd.ic_isotropic(48, 40, 32, k0=6.0, urms=1.0, mt=0.5, seed=4242)
    assert hashlib.sha256(Q.tobytes(order="F")).hexdigest() == out[0]
