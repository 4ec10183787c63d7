"""Pins of the oracle's operator matrices and LU against independent maths.

mat_build / mat_build_IRK / mat_build_uphibar (mat_assembly.F90:68-413):
each matrix is applied to the scaled Chebyshev coefficients of a known
polynomial and must return the continuous operator evaluated at the
collocation points (computed with numpy's Chebyshev algebra, not the oracle),
with the boundary rows giving the boundary values.

dgetrf/dgetrs (algebra_lapack.F90:14-53): reconstruction P A = L U, LAPACK's
first-maximum pivot rule, INFO on a singular matrix, and solutions against
numpy.linalg.solve.
"""
from __future__ import annotations

import numpy as np
import numpy.polynomial.chebyshev as npc
import pytest

from conftest import rel
from test_oracle_transforms import ETA, cheb_points, series_to_ft


def poly_in_r(N, seed):
    """A random polynomial of degree N-1 (decaying coefficients) with its
    first and second r-derivatives on the grid; d/dr = 2 d/dx."""
    rng = np.random.default_rng(seed)
    c = rng.standard_normal(N) / (1.0 + np.arange(N)) ** 3
    x = cheb_points(N)
    f = npc.chebval(x, c)
    f1 = 2 * npc.chebval(x, npc.chebder(c))
    f2 = 4 * npc.chebval(x, npc.chebder(c, 2))
    return series_to_ft(c).real, f, f1, f2


@pytest.mark.parametrize("N,m", [(17, 0), (17, 3), (33, 7), (65, 1)])
def test_AT_applies_helmholtz_with_dirichlet_rows(orc, N, m):
    wdt, Pr = 0.5 * 1e-3, 0.7
    AT, _, _ = orc.mat_build(N, ETA, m, wdt, Pr)
    r, _, _ = orc.grid(N, ETA)
    ft, f, f1, f2 = poly_in_r(N, N + m)
    want = f - (wdt / Pr) * (f2 + f1 / r - m ** 2 * f / r ** 2)
    want[0], want[-1] = f[0], f[-1]
    got = AT @ ft
    assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()


@pytest.mark.parametrize("irk", [False, True])
@pytest.mark.parametrize("mBC", ["NS", "SF"])
@pytest.mark.parametrize("N,m", [(17, 1), (33, 5)])
def test_AF_applies_streamfunction_vorticity_system(orc, N, m, mBC, irk):
    wdt = 0.5 * 1e-3
    _, AF, _ = orc.mat_build(N, ETA, m, wdt, 1.0, mBC=mBC, irk_order=irk)
    r, _, _ = orc.grid(N, ETA)
    fp, p, p1, p2 = poly_in_r(N, 2 * N + m)
    fw, w, w1, w2 = poly_in_r(N, 3 * N + m)
    lap_p = p2 + p1 / r - m ** 2 * p / r ** 2
    lap_w = w2 + w1 / r - m ** 2 * w / r ** 2
    top = lap_p + w                      # rows 1..N:   L_m psi + omega = 0
    bot = w - wdt * lap_w                # rows N+1..2N: (1 - w dt L_m) omega
    bc_psi = (p[0], p[-1])
    bc_d = (p1[0], p1[-1]) if mBC == "NS" else (p2[0] - p1[0] / r[0], p2[-1] - p1[-1] / r[-1])
    first, second = (bc_d, bc_psi) if irk else (bc_psi, bc_d)
    top[0], top[-1] = first
    bot[0], bot[-1] = second
    got = AF @ np.concatenate([fp, fw])
    want = np.concatenate([top, bot])
    assert np.abs(got - want).max() <= 1e-11 * np.abs(want).max()


@pytest.mark.parametrize("mBC", ["NS", "SF"])
def test_Auphi_mean_flow_operator(orc, mBC):
    N, wdt = 33, 0.5 * 1e-3
    _, _, Au = orc.mat_build(N, ETA, 0, wdt, 1.0, mBC=mBC)
    r, _, _ = orc.grid(N, ETA)
    ft, f, f1, f2 = poly_in_r(N, 77)
    want = f - wdt * (f2 + f1 / r - f / r ** 2)
    if mBC == "NS":
        want[0], want[-1] = f[0], f[-1]
    else:
        want[0], want[-1] = f1[0] - f[0] / r[0], f1[-1] - f[-1] / r[-1]
    assert np.abs(Au @ ft - want).max() <= 1e-12 * np.abs(want).max()


def test_AF_independent_of_Pr_and_AT_depends_on_it(orc):
    a1 = orc.mat_build(17, ETA, 3, 5e-4, 1.0)
    a2 = orc.mat_build(17, ETA, 3, 5e-4, 7.0)
    assert np.array_equal(a1[1], a2[1]) and np.array_equal(a1[2], a2[2])
    assert not np.array_equal(a1[0], a2[0])


# --------------------------------------------------------------------- LU
def unpack(LU, piv):
    n = LU.shape[0]
    L = np.tril(LU, -1) + np.eye(n)
    U = np.triu(LU)
    P = np.eye(n)
    for j, p in enumerate(piv):
        P[[j, p]] = P[[p, j]]
    return P, L, U


@pytest.mark.parametrize("n", [1, 2, 7, 40, 130])
def test_dgetrf_reconstructs(orc, n):
    rng = np.random.default_rng(n)
    A = rng.standard_normal((n, n))
    LU, piv, info = orc.dgetrf(A)
    assert info == 0
    P, L, U = unpack(LU, piv)
    assert np.abs(P @ A - L @ U).max() < 1e-13 * n
    # partial pivoting: |l_ij| <= 1
    assert np.abs(np.tril(LU, -1)).max() <= 1.0


def test_dgetrf_first_max_pivot_and_info(orc):
    A = np.array([[1.0, 2.0, 0.0], [-3.0, 1.0, 1.0], [3.0, 0.0, 1.0]])
    LU, piv, info = orc.dgetrf(A)
    assert piv[0] == 1          # |-3| and |3| tie: the first one wins (idamax)
    S = np.array([[1.0, 2.0], [2.0, 4.0]])
    _, _, info = orc.dgetrf(S)
    assert info == 2            # U(2,2) is exactly zero, INFO = 2


@pytest.mark.parametrize("which", ["AT", "AF", "Au"])
@pytest.mark.parametrize("N,m", [(17, 2), (65, 40)])
def test_dgetrs_solves_operator_systems(orc, which, N, m):
    mats = dict(zip(("AT", "AF", "Au"), orc.mat_build(N, ETA, m, 0.5e-5, 1.0)))
    A = mats[which]
    rng = np.random.default_rng(N + m)
    b = rng.standard_normal(A.shape[0])
    LU, piv, info = orc.dgetrf(A)
    assert info == 0
    x = orc.dgetrs(LU, piv, b)
    ref = np.linalg.solve(A, b)
    assert rel(x, ref) < 1e-10
    assert np.abs(A @ x - b).max() < 1e-12 * np.abs(b).max() * np.abs(A).max()
