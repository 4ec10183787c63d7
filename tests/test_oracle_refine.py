"""The oracle's restart with a change of resolution (init.F90:722-861):
REDFT00 of each mode on the checkpoint's grid, zero-padded, REDFT00 on the
new grid over 2 (Nr - 1), times 2 (Nr - 1) / (2 (Nr_old - 1)).

Pinned by what the mathematics fixes: REDFT00 applied twice is 2 (n - 1)
times the identity, so an unchanged radial grid keeps every value; on a finer
grid, a polynomial of degree below Nr_old - 1 is the Chebyshev interpolant
evaluated at the new points (numpy), while the highest old mode T_{Nr_old-1}
comes out doubled (the last DCT-I term's halving is not undone there); new
azimuthal modes start at zero; a coarser grid is refused.
"""
from __future__ import annotations

import numpy as np
import numpy.polynomial.chebyshev as npc
import pytest

from test_oracle_transforms import cheb_points


def _state(Nm, Nr, seed):
    rng = np.random.default_rng(seed)
    return tuple(np.asfortranarray(rng.standard_normal((Nm + 1, Nr)) + 1j * rng.standard_normal((Nm + 1, Nr)))
                 for _ in range(4))


def test_same_size_copies(orc):
    st = _state(4, 9, 1)
    out = orc.refine(st, 4, 9)
    for a, b in zip(st, out):
        assert np.array_equal(a, b)


def test_more_modes_same_grid(orc):
    st = _state(4, 9, 2)
    out = orc.refine(st, 7, 9)
    for a, b in zip(st, out):
        assert np.abs(b[:5] - a).max() < 1e-13 * np.abs(a).max()
        assert not b[5:].any()


@pytest.mark.parametrize("Nm_new", [3, 6])
def test_finer_grid_interpolates(orc, Nm_new):
    No, Nn, Nm = 9, 17, 3
    xo, xn = cheb_points(No), cheb_points(Nn)
    rng = np.random.default_rng(3)
    # degree No - 2 per mode: the last Chebyshev coefficient is zero
    cre = rng.standard_normal((Nm + 1, No - 1))
    cim = rng.standard_normal((Nm + 1, No - 1))
    f = np.array([npc.chebval(xo, cre[m]) + 1j * npc.chebval(xo, cim[m]) for m in range(Nm + 1)])
    want = np.array([npc.chebval(xn, cre[m]) + 1j * npc.chebval(xn, cim[m]) for m in range(Nm + 1)])
    out = orc.refine((f, 2 * f, -f, 1j * f), Nm_new, Nn)
    assert np.abs(out[0][:Nm + 1] - want).max() < 1e-12 * np.abs(want).max()
    assert np.abs(out[3][:Nm + 1] - 1j * want).max() < 1e-12 * np.abs(want).max()
    assert not out[0][Nm + 1:].any()


def test_highest_old_mode_doubled(orc):
    No, Nn = 9, 13
    xo, xn = cheb_points(No), cheb_points(Nn)
    e = np.zeros(No)
    e[-1] = 1.0
    f = npc.chebval(xo, e)[None, :].astype(complex)
    out = orc.refine((f, f, f, f), 0, Nn)
    assert np.abs(out[0][0] - 2 * npc.chebval(xn, e)).max() < 1e-12


def test_coarser_grid_refused(orc):
    with pytest.raises(ValueError):
        orc.refine(_state(4, 9, 4), 4, 5)
    with pytest.raises(ValueError):
        orc.refine(_state(4, 9, 4), 2, 9)
