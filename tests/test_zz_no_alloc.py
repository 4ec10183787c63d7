"""fdns_budget_noalloc, the allocation-free variant of fdns_budget, through the
C-ABI: two calls in a row, the second on a run with fewer radial points, each
equal to fdns_budget bit for bit.  Named to be collected last.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _budget(lib, h, N, fn):
    out4 = np.zeros(4)
    upbal = np.zeros(3 * N)
    rc = getattr(lib, fn)(h, out4.ctypes.data_as(C.c_void_p), upbal.ctypes.data_as(C.c_void_p))
    assert rc == 0, lib.fdns_strerror(rc)
    return out4, upbal


def test_budget_noalloc_matches_budget(fd, orc):
    lib = fd.lib()
    lib.fdns_budget_noalloc.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
    lib.fdns_budget_noalloc.restype = C.c_int
    runs = []
    for N, Nm, scheme in ((65, 32, "ARS443"), (33, 16, "CNAB2")):     # the second call is the smaller
        r = fd.Run(N, Nm, 1e5, 1.0, 1e-5, scheme, n_KE=1)
        r.set_state(*orc.seed_state(N, Nm, 0.35, 7, amp=0.3))
        r.step(3)
        runs.append((r, N))
    for r, N in runs:
        a4, au = _budget(lib, r.h, N, "fdns_budget")
        b4, bu = _budget(lib, r.h, N, "fdns_budget_noalloc")
        assert np.array_equal(a4, b4) and np.array_equal(au, bu)
        assert np.abs(au).max() > 0 and a4[3] > 0
    assert lib.fdns_budget_noalloc(runs[0][0].h, None, None) == -1      # null output refused
    for r, _ in runs:
        r.close()
