"""CPU-side checks of the C-ABI library: it loads, and it exports every entry
point include/sfem_capi.h declares (no compute calls without a GPU)."""
import os
import re

import pytest

from solidfem_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", fn)).read()
            names |= set(re.findall(r"^\s*(?:int|void|const char\*)\s+(sfem_\w+)\s*\(", txt, re.M))
    return names


def test_library_exports_every_declared_symbol():
    L = capi.lib()
    names = declared()
    assert len(names) >= 17
    for n in sorted(names):
        assert hasattr(L, n), n
    assert set(capi.EXPORTS) == names


def test_create_fails_cleanly_without_device_or_bad_input():
    import numpy as np
    X = np.zeros((4, 3))
    X[1, 0] = X[2, 1] = X[3, 2] = 1.0
    # out-of-range node index is rejected before any device call
    with pytest.raises(capi.SfemError) as e:
        capi.Solver(X, [[0, 1, 2, 7]], [0, 0, 0, 0], capi.LINEAR)
    assert "out of range" in str(e.value)
    with pytest.raises(capi.SfemError):
        capi.Solver(X, [[0, 1, 2, 3]], [0, 0, 0, 0], 9)
    err = capi.lib().sfem_last_error()
    assert isinstance(err, bytes) and len(err) > 0
