"""Shared test helpers: seeded smooth states, closed-form symbols, error norms.

Arrays follow the solver layout F(My, Mz, Mx) in Fortran order, so axis 0
is y, axis 1 is z, axis 2 is x.
"""
import numpy as np

AXIS = {0: 2, 1: 0, 2: 1}   # direction (x, y, z) -> numpy axis


def grid_1d(n, period=2 * np.pi):
    return np.arange(n) * period / n


def along(d, g, shape):
    """Broadcast a 1-D profile g along direction d into a field of `shape`."""
    s = [1, 1, 1]
    s[AXIS[d]] = len(g)
    return np.asfortranarray(np.broadcast_to(g.reshape(s), shape).copy())


def coords(nx, ny, nz, period=2 * np.pi):
    x, y, z = grid_1d(nx, period), grid_1d(ny, period), grid_1d(nz, period)
    X = x[None, None, :]
    Y = y[:, None, None]
    Z = z[None, :, None]
    return X, Y, Z


def smooth_state(nx, ny, nz, seed=0, amp=1.0, mt=0.3, gamma=1.4, nmodes=6):
    """A smooth, seeded, periodic compressible state Q(My,Mz,Mx,5)."""
    rng = np.random.default_rng(seed)
    X, Y, Z = coords(nx, ny, nz)
    shape = (ny, nz, nx)

    def field(scale):
        f = np.zeros(shape)
        for _ in range(nmodes):
            k = rng.integers(-3, 4, size=3)
            ph = rng.uniform(0, 2 * np.pi)
            a = rng.normal() * scale
            f = f + a * np.sin(k[0] * X + k[1] * Y + k[2] * Z + ph)
        return f

    u, v, w = field(amp), field(amp), field(amp)
    rho = 1.0 + 0.05 * np.tanh(field(1.0))
    p = 1.0 / (gamma * mt * mt) * (1.0 + 0.05 * np.tanh(field(1.0)))
    Q = np.zeros(shape + (5,), order="F")
    Q[..., 0] = rho * u
    Q[..., 1] = rho * v
    Q[..., 2] = rho * w
    Q[..., 3] = rho
    Q[..., 4] = p / (gamma - 1.0) + 0.5 * rho * (u * u + v * v + w * w)
    return Q


def relerr(a, b, axis_fields=True):
    """Norm-wise relative error per field: max|a-b| / max|b| (field axis last)."""
    a = np.asarray(a)
    b = np.asarray(b)
    if a.ndim == 4 and axis_fields:
        num = np.abs(a - b).reshape(-1, a.shape[3]).max(axis=0)
        den = np.abs(b).reshape(-1, b.shape[3]).max(axis=0)
        den = np.where(den == 0, 1.0, den)
        return num / den
    den = np.abs(b).max()
    return np.abs(a - b).max() / (den if den else 1.0)


# ---- closed-form symbols (theta = k h) ----
def sym_d1(th):
    return ((14.0 / 9.0) * np.sin(th) + (1.0 / 18.0) * np.sin(2 * th)) / (1 + (2.0 / 3.0) * np.cos(th))


def sym_d2(th, shipped=False):
    if shipped:
        a2, b2 = 12.0 / 11.0, 3.0 / 44.0
        c2 = 2 * (a2 + b2)
        num = 2 * a2 * np.cos(th) + 2 * b2 * np.cos(2 * th) + 2 * c2 * np.cos(3 * th)
        return -num / (1 + (4.0 / 11.0) * np.cos(th))   # returned as "k~^2 h^2" with D2 f = -sym f
    return ((24.0 / 11.0) * (1 - np.cos(th)) + (3.0 / 22.0) * (1 - np.cos(2 * th))) / (1 + (4.0 / 11.0) * np.cos(th))


FILT_A = (0.99965, 0.66652, 0.16674, 4.0e-5, -5.0e-6)
FILT_AL, FILT_BE = 0.66624, 0.16688
GAUSS_A = (3565.0 / 10368.0, 3091.0 / 12960.0, 1997.0 / 25920.0, 149.0 / 12960.0, 107.0 / 103680.0)


def sym_filter(th):
    a = FILT_A
    num = a[0] + 2 * sum(a[m] * np.cos(m * th) for m in range(1, 5))
    return num / (1 + 2 * FILT_AL * np.cos(th) + 2 * FILT_BE * np.cos(2 * th))


def sym_gauss(th):
    a = GAUSS_A
    return a[0] + 2 * sum(a[m] * np.cos(m * th) for m in range(1, 5))


def spectral_deriv(f, axis, order=1, period=2 * np.pi):
    n = f.shape[axis]
    k = np.fft.fftfreq(n, d=period / n) * 2 * np.pi
    if order == 1:
        k = k.copy()
        if n % 2 == 0:
            k[n // 2] = 0.0
        mult = 1j * k
    else:
        mult = -(k ** 2)
    sh = [1] * f.ndim
    sh[axis] = n
    return np.real(np.fft.ifft(np.fft.fft(f, axis=axis) * mult.reshape(sh), axis=axis))
