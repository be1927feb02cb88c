"""The fused stage update: the last (x) pass of the artificial-transport
divergence also applies the sponge term (when set), the low-storage RK
update and the x filter.  The fused and the separate passes
(ST_NO_FUSE_RK=1) round alike, so a march gives the same bits either way; both are checked against the oracle by the
parity tests."""
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

st_amd = pytest.importorskip("st_amd")


def _log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


PER = (2.0, 2 * np.pi, 2 * np.pi)


def _march(shape, nsteps, fuse, sponge=False, **kw):
    if sponge:
        Q, s1, s2 = st_amd.ic_shock(*shape, PER, mach=1.5, frac=0.25, urms=0.1, seed=3)
    else:
        Q = st_amd.ic_isotropic(*shape, k0=4.0, urms=1.0, mt=0.3, seed=77)
    if not fuse:
        os.environ["ST_NO_FUSE_RK"] = "1"
    # one lane layout for every pass: the fused and the separate divergence
    # passes have launch shapes of their own, and the layout sets the
    # chunking of the line recurrences, hence the rounding
    os.environ["ST_NARROW"] = "0"
    try:
        s = st_amd.Solver(*shape, **kw)
        try:
            if sponge:
                s.sponge(5.0, 2.0, 0.4, s1, s2)
            s.upload(Q)
            t, st, dts = s.march(nsteps)
            return s.download(), t, dts
        finally:
            s.close()
    finally:
        os.environ.pop("ST_NO_FUSE_RK", None)
        os.environ.pop("ST_NARROW", None)


@pytest.mark.parametrize("shape,kw", [((32, 32, 32), {}), ((64, 16, 32), {}),
                                       ((64, 16, 16), {"nonperiodic_x": 1, "period": PER}),
                                       ((64, 16, 16), {"nonperiodic_x": 1, "period": PER, "sponge": True})])
def test_fused_update_same_bits(shape, kw):
    """With the sponge set, its term is folded into the fused update too."""
    _log("fused vs separate update", shape, kw)
    a = _march(shape, 3, True, Re=150.0, hyper_param=1, **kw)
    b = _march(shape, 3, False, Re=150.0, hyper_param=1, **kw)
    assert a[1] == b[1] and np.array_equal(a[2], b[2])
    assert np.array_equal(a[0], b[0])
