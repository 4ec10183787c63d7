"""Shared set-up of the test suite.

Marker ``gpu``: the test drives libfdns.so on a HIP device.  Everything else
runs on the CPU (the oracle, the host logic of the C-ABI, the golden
fixtures).
"""
from __future__ import annotations

import os
import sys

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(REPO, "falcon-dns_amd"))
sys.path.insert(0, os.path.join(REPO, "oracle"))
GOLDEN = os.path.join(REPO, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X) and the built libfdns.so")


@pytest.fixture(scope="session")
def orc():
    """The oracle (CPU restatement), built on first use."""
    import oracle
    if not os.path.exists(oracle._LIB_PATH):
        oracle.build()
    return oracle


@pytest.fixture(scope="session")
def fd():
    """The product library; on a GPU run it must be present and load."""
    import fdns
    fdns.lib()
    return fdns


def rel(a, b):
    """Relative Frobenius error ||a - b|| / ||b|| (absolute when b == 0)."""
    a = np.asarray(a)
    b = np.asarray(b)
    n = np.linalg.norm(b)
    d = np.linalg.norm(a - b)
    return d / n if n > 0 else d
