"""st_cfl_noalloc (the Compute_CFL entry point that allocates nothing)
against st_cfl on the same states, through the C-ABI: two calls in a row,
the second on a smaller grid, each bit-identical to st_cfl (maxima do not
depend on the reduction order); a grid above the scratch maximum is
refused with a message, before any launch."""
import ctypes
import os

import numpy as np
import pytest

import st_amd
from fields import smooth_state

pytestmark = pytest.mark.gpu


def _hip():
    for name in ("/opt/rocm/lib/libamdhip64.so", "libamdhip64.so"):
        try:
            h = ctypes.CDLL(name)
            break
        except OSError:
            continue
    else:
        pytest.skip("libamdhip64 not found")
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipFree.argtypes = [ctypes.c_void_p]
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    h.hipDeviceSynchronize.argtypes = []
    return h


def _noalloc(L, grid, gamma, q_dev, out_dev):
    return L.st_cfl_noalloc(ctypes.byref(grid), gamma, q_dev, out_dev, None)


def test_cfl_noalloc_matches_st_cfl():
    L = st_amd.lib()
    hip = _hip()
    out_dev = ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(out_dev), 16) == 0           # the test's own output buffer
    try:
        for shape, seed in (((64, 64, 64), 3), ((32, 48, 16), 4)):   # the second call is smaller
            print("cfl_noalloc", shape, flush=True)
            s = st_amd.Solver(*shape, Re=100.0)
            try:
                Q = smooth_state(*shape, seed=seed, amp=1.3, mt=0.4)
                s.upload(Q)
                stab, resv = s.cfl()
                qd = ctypes.POINTER(ctypes.c_double)()
                assert L.st_state_dev(s.h, ctypes.byref(qd)) == 0
                rc = _noalloc(L, s.grid, 1.4, ctypes.cast(qd, ctypes.c_void_p), out_dev)
                assert rc == 0, L.st_last_error(None)
                assert hip.hipDeviceSynchronize() == 0
                got = np.zeros(2)
                assert hip.hipMemcpy(got.ctypes.data_as(ctypes.c_void_p), out_dev, 16, 2) == 0   # device to host
                assert got[0] == stab and got[1] == resv, (got, stab, resv)
                # sizes above the scratch maximum are refused before any launch
                big = st_amd.Grid(512, 512, 256, 1.0, 1.0, 1.0, 0.0, 0.0, 0.0, 0)
                rc = _noalloc(L, big, 1.4, ctypes.cast(qd, ctypes.c_void_p), out_dev)
                assert rc == -2 and b"exceed the maximum" in L.st_last_error(None)
            finally:
                s.close()
    finally:
        hip.hipFree(out_dev)
