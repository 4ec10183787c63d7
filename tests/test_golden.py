"""The oracle reproduces the committed golden fixtures (tests/golden/*.npz).

The fixtures were written by tests/golden/make_golden.py from the oracle;
this keeps the oracle and the fixtures in step, so that the GPU tests, which
compare against both, compare against the same thing.
"""
from __future__ import annotations

import ast
import glob
import os

import numpy as np
import pytest

from conftest import GOLDEN, rel

FILES = sorted(glob.glob(os.path.join(GOLDEN, "*.npz")))


def load(path):
    z = np.load(path)            # plain arrays only, no pickled objects
    case = ast.literal_eval(str(z["params"]))
    return case, {k: z[k] for k in ("tFR", "psi", "series")}


def test_fixtures_present():
    names = {os.path.basename(f) for f in FILES}
    assert "c0_cnab2_nr65_nm128.npz" in names and len(names) >= 10


@pytest.mark.parametrize("path", FILES, ids=[os.path.basename(f)[:-4] for f in FILES])
def test_oracle_reproduces_fixture(orc, path):
    import sys
    sys.path.insert(0, GOLDEN)
    from make_golden import run_case
    case, want = load(path)
    got = run_case(*case)
    for k in ("tFR", "psi"):
        assert rel(got[k], want[k]) < 1e-12, k
    assert np.abs(got["series"] - want["series"]).max() <= 1e-12 * np.abs(want["series"]).max()
    # sanity of the physics in the fixture: finite, heat flux near conduction
    # (except the zero-history BDF3 start, whose first step scales the
    # interior temperature by 18/11, timeschemes.F90:111-113)
    assert np.all(np.isfinite(want["series"]))
    if "fresh" not in path:
        assert np.all(np.abs(want["series"][:, 2] - 1) < 0.5)
