"""CPU checks of the C-ABI boundary: the shared library loads and exports
every entry point that include/st_api.h declares (no GPU calls)."""
import os
import re

import st_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    txt = open(os.path.join(ROOT, "include", "st_api.h")).read()
    return sorted(set(re.findall(r"\b(st_[a-z0-9_]+)\s*\(", txt)))


def test_exports():
    lib = st_amd.lib()
    names = declared()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert sorted(st_amd.EXPORTS) == names


def test_ic_host_side():
    import numpy as np
    Q = st_amd.ic_isotropic(16, 16, 16, k0=3.0, urms=1.0, mt=0.3, seed=3)
    assert Q.shape == (16, 16, 16, 5)
    assert np.isfinite(Q).all()


def test_library_built_from_these_sources():
    """libst.so carries the hash of the sources it was built from; it must
    be the hash of the sources in this tree (a stale library fails here)."""
    assert st_amd.build_id() == st_amd.source_id(), (st_amd.build_id(), st_amd.source_id())
