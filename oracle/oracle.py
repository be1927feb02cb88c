"""ctypes view of the CPU oracle (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
import this module.  Arrays are numpy float64 in Fortran order with
shape (My, Mz, Mx) per field, (My, Mz, Mx, nf) for several fields.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")

_dp = ctypes.POINTER(ctypes.c_double)


class Params(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
        ("period_x", ctypes.c_double), ("period_y", ctypes.c_double), ("period_z", ctypes.c_double),
        ("x_start", ctypes.c_double), ("y_start", ctypes.c_double), ("z_start", ctypes.c_double),
        ("Re", ctypes.c_double), ("Pr", ctypes.c_double), ("gamma", ctypes.c_double),
        ("vis_exp", ctypes.c_double), ("bulk_ratio", ctypes.c_double),
        ("hyper_param", ctypes.c_int),
        ("cfl_cal", ctypes.c_double),
        ("d2_mode", ctypes.c_int), ("rk_scheme", ctypes.c_int), ("mutate_q", ctypes.c_int),
        ("nonper_x", ctypes.c_int),
    ]


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _load():
    if not os.path.exists(_LIB):
        build()
    lib = ctypes.CDLL(_LIB)
    lib.or_create.restype = ctypes.c_void_p
    lib.or_create.argtypes = [ctypes.POINTER(Params)]
    lib.or_destroy.argtypes = [ctypes.c_void_p]
    lib.or_threads.argtypes = [ctypes.c_int]
    lib.or_threads.restype = ctypes.c_int
    lib.or_metrics.argtypes = [ctypes.c_void_p, _dp]
    lib.or_dfd.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp, ctypes.c_int, ctypes.c_double]
    lib.or_d2fd2.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp, _dp, ctypes.c_int]
    lib.or_compact_filter.argtypes = [ctypes.c_void_p, _dp, _dp, ctypes.c_int]
    lib.or_gaussian_filter.argtypes = [ctypes.c_void_p, _dp, _dp, ctypes.c_int]
    lib.or_filter_dir.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp]
    lib.or_gauss_dir.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp]
    lib.or_band_solve.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, _dp, ctypes.c_long]
    lib.or_np_lhs.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp]
    lib.or_np_band_solve.argtypes = [ctypes.c_int, _dp, _dp, ctypes.c_long]
    lib.or_set_sponge.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double, ctypes.c_double, _dp, _dp]
    lib.or_rhs.argtypes = [ctypes.c_void_p, _dp, _dp]
    lib.or_compute_hyper.argtypes = [ctypes.c_void_p, _dp]
    lib.or_hyper_rhs.argtypes = [ctypes.c_void_p, _dp, _dp]
    lib.or_get_grad.argtypes = [ctypes.c_void_p, _dp]
    lib.or_get_hyper.argtypes = [ctypes.c_void_p, _dp]
    lib.or_set_hyper.argtypes = [ctypes.c_void_p, _dp]
    lib.or_fix_pressure.argtypes = [ctypes.c_void_p, _dp]
    lib.or_compute_cfl.argtypes = [ctypes.c_void_p, _dp, _dp, _dp]
    lib.or_compute_tstep.argtypes = [ctypes.c_void_p, _dp, ctypes.c_double, ctypes.c_int]
    lib.or_compute_tstep.restype = ctypes.c_double
    lib.or_rk_stage.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, ctypes.c_double, ctypes.c_double, ctypes.c_double]
    lib.or_lddrk4.argtypes = [ctypes.c_void_p, _dp, ctypes.c_double, _dp, ctypes.POINTER(ctypes.c_int)]
    lib.or_march.argtypes = [ctypes.c_void_p, _dp, _dp, ctypes.POINTER(ctypes.c_int)]
    lib.or_march.restype = ctypes.c_double
    lib.or_diagnostics.argtypes = [ctypes.c_void_p, _dp, _dp]
    lib.or_compute_grad.argtypes = [ctypes.c_void_p, _dp, _dp, _dp]
    lib.or_clamps.argtypes = [ctypes.c_void_p]
    lib.or_clamps.restype = ctypes.c_long
    lib.or_rk_coeffs.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp]
    lib.or_rk_coeffs.restype = ctypes.c_int
    return lib


_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = _load()
    return _lib


def _p(a):
    assert a.dtype == np.float64 and (a.flags.f_contiguous or a.ndim == 1)
    return a.ctypes.data_as(_dp)


def fzeros(shape):
    return np.zeros(shape, dtype=np.float64, order="F")


def band_solve(R, alpha, beta):
    """Cyclic (beta, alpha, 1, alpha, beta) solve of each row of R (C-order rows)."""
    R = np.ascontiguousarray(R, dtype=np.float64).copy()
    n = R.shape[-1]
    lines = R.size // n
    lib().or_band_solve(n, alpha, beta, R.ctypes.data_as(_dp), lines)
    return R


def np_band_solve(rows, R):
    """pentadiagonal_serial with band rows (n, 5) = (A, B, C, D, E) on each row of R."""
    rows = np.ascontiguousarray(rows, dtype=np.float64)
    R = np.ascontiguousarray(R, dtype=np.float64).copy()
    n = R.shape[-1]
    assert rows.shape == (n, 5)
    lib().or_np_band_solve(n, rows.ctypes.data_as(_dp), R.ctypes.data_as(_dp), R.size // n)
    return R


def rk_coeffs(scheme, loop):
    a = np.zeros(6)
    b = np.zeros(6)
    n = lib().or_rk_coeffs(scheme, loop, _p(a), _p(b))
    return a[:n], b[:n]


class Oracle:
    """One oracle context for a fixed grid and parameter set."""

    def __init__(self, nx, ny, nz, period=(2 * np.pi,) * 3, Re=100.0, Pr=0.7, gamma=1.4,
                 vis_exp=0.76, bulk_ratio=0.0, hyper_param=1, cfl_cal=0.5,
                 d2_mode=0, rk_scheme=0, mutate_q=0, start=(0.0, 0.0, 0.0), nonper_x=0):
        self.params = Params(nx, ny, nz, period[0], period[1], period[2], start[0], start[1], start[2],
                             Re, Pr, gamma, vis_exp, bulk_ratio, hyper_param, cfl_cal,
                             d2_mode, rk_scheme, mutate_q, nonper_x)
        self.shape = (ny, nz, nx)
        self.h = lib().or_create(ctypes.byref(self.params))

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_destroy(self.h)
            self.h = None

    def np_lhs(self, op):
        """LHS band rows (n, 5) of the non-periodic x operator op (0 D1, 1 D2, 2 filter)."""
        r = np.zeros((self.params.nx, 5))
        lib().or_np_lhs(self.h, op, r.ctypes.data_as(_dp))
        return r

    def set_sponge(self, amp, power, frac, ref_left, ref_right):
        """Sponge_Forcing after Hyper_RHS in every stage (amp 0: off)."""
        L = np.ascontiguousarray(ref_left, dtype=np.float64)
        R = np.ascontiguousarray(ref_right, dtype=np.float64)
        lib().or_set_sponge(self.h, amp, power, frac, L.ctypes.data_as(_dp), R.ctypes.data_as(_dp))

    def metrics(self):
        out = np.zeros(9)
        lib().or_metrics(self.h, _p(out))
        return {"sx1": out[0:3], "sx2": out[3:6], "D": out[6:9]}

    def _f(self, nf=None):
        return fzeros(self.shape if nf is None else self.shape + (nf,))

    def dfd(self, d, F, dF=None, iadd=0, cadd=1.0):
        F = np.asfortranarray(F)
        out = self._f() if dF is None else np.asfortranarray(dF).copy(order="F")
        lib().or_dfd(self.h, d, _p(F), _p(out), iadd, cadd)
        return out

    def d2fd2(self, d, F, iadd=0):
        F = np.asfortranarray(F)
        dF, d2F = self._f(), self._f()
        lib().or_d2fd2(self.h, d, _p(F), _p(dF), _p(d2F), iadd)
        return (dF, d2F) if iadd else d2F

    def filter_dir(self, d, F):
        F = np.asfortranarray(F)
        out = self._f()
        lib().or_filter_dir(self.h, d, _p(F), _p(out))
        return out

    def gauss_dir(self, d, F):
        F = np.asfortranarray(F)
        out = self._f()
        lib().or_gauss_dir(self.h, d, _p(F), _p(out))
        return out

    def compact_filter(self, F):
        F = np.asfortranarray(F)
        nf = F.shape[3] if F.ndim == 4 else 1
        out = np.zeros_like(F, order="F")
        lib().or_compact_filter(self.h, _p(F), _p(out), nf)
        return out

    def gaussian_filter(self, F):
        F = np.asfortranarray(F)
        nf = F.shape[3] if F.ndim == 4 else 1
        out = np.zeros_like(F, order="F")
        lib().or_gaussian_filter(self.h, _p(F), _p(out), nf)
        return out

    def rhs(self, Q):
        """Navier-Stokes RHS; returns (Qrhs, Q_after) (Q_after differs only if mutate_q)."""
        Qw = np.array(Q, dtype=np.float64, order="F", copy=True)
        out = self._f(5)
        lib().or_rhs(self.h, _p(Qw), _p(out))
        return out, Qw

    def hyper_rhs(self, Q, Qrhs):
        Q = np.asfortranarray(Q)
        out = np.array(Qrhs, dtype=np.float64, order="F", copy=True)
        lib().or_hyper_rhs(self.h, _p(Q), _p(out))
        return out

    def full_rhs(self, Q):
        r, Qa = self.rhs(Q)
        return self.hyper_rhs(Qa, r)

    def grad(self):
        out = self._f(12)
        lib().or_get_grad(self.h, _p(out))
        return out

    def hyper(self):
        out = self._f(3)
        lib().or_get_hyper(self.h, _p(out))
        return out

    def set_hyper(self, H):
        H = np.asfortranarray(H)
        lib().or_set_hyper(self.h, _p(H))

    def compute_hyper(self, Q):
        Q = np.asfortranarray(Q)
        lib().or_compute_hyper(self.h, _p(Q))
        return self.hyper()

    def fix_pressure(self, Q):
        Qw = np.array(Q, dtype=np.float64, order="F", copy=True)
        lib().or_fix_pressure(self.h, _p(Qw))
        return Qw

    def cfl(self, Q):
        Q = np.asfortranarray(Q)
        a = np.zeros(1)
        b = np.zeros(1)
        lib().or_compute_cfl(self.h, _p(Q), _p(a), _p(b))
        return a[0], b[0]

    def tstep(self, Q, cfl_resv, step):
        Q = np.asfortranarray(Q)
        return lib().or_compute_tstep(self.h, _p(Q), cfl_resv, step)

    def rk_stage(self, Q, QW, alpha, beta, dt):
        Qw = np.array(Q, dtype=np.float64, order="F", copy=True)
        QWw = np.array(QW, dtype=np.float64, order="F", copy=True)
        dQ = self._f(5)
        lib().or_rk_stage(self.h, _p(Qw), _p(QWw), _p(dQ), alpha, beta, dt)
        return Qw, QWw, dQ

    def lddrk4(self, Q, dt, time=0.0, step=0):
        Qw = np.array(Q, dtype=np.float64, order="F", copy=True)
        t = np.array([time])
        s = ctypes.c_int(step)
        lib().or_lddrk4(self.h, _p(Qw), dt, _p(t), ctypes.byref(s))
        return Qw, t[0], s.value

    def march(self, Q, nsteps, time=0.0, step=0, log=None):
        """Run nsteps of main.f's march mode 1; returns (Q, time, step, dts)."""
        Qw = np.array(Q, dtype=np.float64, order="F", copy=True)
        t = np.array([time])
        s = ctypes.c_int(step)
        dts = []
        for _ in range(nsteps):
            dts.append(lib().or_march(self.h, _p(Qw), _p(t), ctypes.byref(s)))
            if log is not None:
                log.append(self.diagnostics(Qw))
        return Qw, t[0], s.value, dts

    def clamps(self):
        """Points Fix_Pressure has clamped (P < 0 -> base_pressure) on this context."""
        return lib().or_clamps(self.h)

    def compute_grad(self, Q):
        """Compute_Grad (src/grad.f:13-115): (MaxDil, Enstrophy); /GRAD/ velocity part updated."""
        Q = np.asfortranarray(Q)
        m, e = np.zeros(1), np.zeros(1)
        lib().or_compute_grad(self.h, _p(Q), _p(m), _p(e))
        return m[0], e[0]

    def diagnostics(self, Q):
        Q = np.asfortranarray(Q)
        out = np.zeros(7)
        lib().or_diagnostics(self.h, _p(Q), _p(out))
        return out
