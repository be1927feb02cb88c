"""The series power used for the viscosity law (csrc/st_math.h) against the
C library's pow, log and exp, on the host: the header is plain C++ and is
compiled here exactly as the kernels use it."""
import ctypes
import os
import subprocess
import tempfile

import numpy as np
import pytest

HDR = os.path.join(os.path.dirname(__file__), "..", "shock-turbulence_amd", "csrc", "st_math.h")

SRC = r"""
#include "%s"
extern "C" void spow(const double *x, long n, double a, double *y) { for (long i = 0; i < n; i++) y[i] = st::series_pow(x[i], a); }
extern "C" void slog(const double *x, long n, double *y) { for (long i = 0; i < n; i++) y[i] = st::series_log(x[i]); }
extern "C" void sexp(const double *x, long n, double *y) { for (long i = 0; i < n; i++) y[i] = st::series_exp(x[i]); }
"""


@pytest.fixture(scope="module")
def lib():
    d = tempfile.mkdtemp()
    src = os.path.join(d, "m.cpp")
    so = os.path.join(d, "m.so")
    with open(src, "w") as f:
        f.write(SRC % os.path.abspath(HDR))
    subprocess.run(["g++", "-O2", "-ffp-contract=off", "-shared", "-fPIC", src, "-o", so], check=True)
    L = ctypes.CDLL(so)
    P = ctypes.POINTER(ctypes.c_double)
    L.spow.argtypes = [P, ctypes.c_long, ctypes.c_double, P]
    L.slog.argtypes = [P, ctypes.c_long, P]
    L.sexp.argtypes = [P, ctypes.c_long, P]
    return L


def _call(f, x, *extra):
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.empty_like(x)
    P = ctypes.POINTER(ctypes.c_double)
    f(x.ctypes.data_as(P), x.size, *extra, y.ctypes.data_as(P))
    return y


def test_pow_against_libm(lib):
    rng = np.random.default_rng(0)
    # gamma-1 times the temperature: about 1/M_t^2 in these runs, here 2^-8 .. 2^8
    x = np.exp2(rng.uniform(-8, 8, 200000))
    for a in (0.76, 0.5, -0.24, 1.5, 2.0):
        y = _call(lib.spow, x, ctypes.c_double(a))
        ref = np.power(x, a)
        rel = np.abs(y - ref) / ref
        assert rel.max() < 4e-15, (a, rel.max())
        assert np.median(rel) < 3e-16


def test_log_exp_pieces(lib):
    rng = np.random.default_rng(1)
    x = np.exp2(rng.uniform(-30, 30, 100000))
    assert np.abs(_call(lib.slog, x) - np.log(x)).max() < 8e-15
    y = rng.uniform(-700, 700, 100000)
    assert (np.abs(_call(lib.sexp, y) - np.exp(y)) / np.exp(y)).max() < 1e-15
    # exact points
    assert _call(lib.slog, np.array([1.0]))[0] == 0.0
    assert _call(lib.sexp, np.array([0.0]))[0] == 1.0
    assert _call(lib.spow, np.array([4.0]), ctypes.c_double(0.5))[0] == pytest.approx(2.0, rel=1e-15)


def test_pow_nonpositive_falls_back(lib):
    y = _call(lib.spow, np.array([0.0, -1.0, np.inf]), ctypes.c_double(0.76))
    assert y[0] == 0.0 and np.isnan(y[1]) and np.isinf(y[2])
