"""The oracle's explicit RK path (time_scheme_type = 'RK': steptime_rk.F90,
fourierloop_rk.F90, rhs_create_rk.F90).

Pinned by what the mathematics fixes:
* the RK2 (midpoint) and classical RK4 tableaux meet their order conditions,
  with the stage sums read below the diagonal only (RK4's a(2,3) = 1/2 of the
  reference's reshape is never read);
* at rest, a temperature eigenmode of the collocation operator is multiplied
  per step by the truncated exponential sum z^k/k!, k <= 2 or 4 (every term,
  diffusion included, is explicit);
* the coupled nonlinear step converges at order 2 (RK2) and 4 (RK4) in every
  field, no slip and stress free;
* conduction is steady;
* the mean-flow right-hand side equals the numpy restatement: the second-order
  radial operator of numpy's Chebyshev fit of u_phi, minus the azimuthal
  means of u_r du_phi/dr and u_r u_phi / r taken in physical space with
  numpy's inverse real FFT.
"""
from __future__ import annotations

import math

import numpy as np
import numpy.polynomial.chebyshev as npc
import pytest

from conftest import rel
from test_oracle_schemes import collocation_L
from test_oracle_transforms import ETA, cheb_points

ORDER = {"RK2": 2, "RK4": 4}


@pytest.mark.parametrize("name", ["RK2", "RK4"])
def test_rk_tableaux(orc, name):
    info = orc.scheme_info("RK", "NONE", name)
    S = ORDER[name]
    assert info["n_imp"] == info["n_exp"] == S
    A, b = info["aA"], info["bA"]
    L = np.tril(A, -1)                    # what the stage sums read
    c = L.sum(1)
    assert abs(b.sum() - 1) < 1e-15 and abs(b @ c - 0.5) < 1e-15
    if name == "RK4":
        assert A[1, 2] == 0.5             # the reference's entry above the diagonal
        for lhs, rhs in ((b @ c ** 2, 1 / 3), (b @ L @ c, 1 / 6), (b @ c ** 3, 1 / 4),
                         (b @ (c * (L @ c)), 1 / 8), (b @ L @ c ** 2, 1 / 12), (b @ L @ L @ c, 1 / 24)):
            assert abs(lhs - rhs) < 1e-15
    else:
        assert abs(b @ c ** 2 - 1 / 3) > 0.05
    with pytest.raises(ValueError):
        orc.scheme_info("RK", "CN", name)


@pytest.mark.parametrize("name", ["RK2", "RK4"])
def test_rk_eigenmode_amplification(orc, name):
    N, Nm, m = 17, 4, 2
    L = collocation_L(N, m)
    lam, V = np.linalg.eig(L)
    k = int(np.argmin(np.abs(lam.real + 2000.0)))   # a mode of |lambda| ~ 2000
    dt = 1.0 / abs(lam[k].real)                     # z = -1
    v = np.zeros(N, complex)
    v[1:-1] = V[:, k] / np.abs(V[:, k]).max() * (0.3 + 0.2j)
    t, o, u, w = orc.seed_state(N, Nm, ETA, 0, amp=0.0)
    t = t.copy()
    t[m, :] = v
    run = orc.Run(N, Nm, 0.0, 1.0, dt, name)
    run.set_state(t, o, u, w)
    z = dt * lam[k].real
    R = sum(z ** j / math.factorial(j) for j in range(ORDER[name] + 1))
    for n in range(1, 4):
        run.step(1)
        assert rel(run.get_state()["tFR"][m], R ** n * v) < 1e-10, n


def _final(orc, scheme, dt, steps, mBC):
    run = orc.Run(17, 8, 2e4, 1.0, dt, scheme, mBC=mBC)
    run.set_state(*orc.seed_state(17, 8, ETA, 5, amp=0.5))
    run.step(steps)
    g = run.get_state()
    return g["tFR"], g["omgFR"], g["upFR"]


@pytest.mark.parametrize("name,mBC", [("RK2", "NS"), ("RK4", "NS"), ("RK4", "SF")])
def test_rk_temporal_order(orc, name, mBC):
    T_end, dt0 = 2e-3, 1e-4
    n0 = round(T_end / dt0)
    ref = _final(orc, name, dt0 / 16, 16 * n0, mBC)
    errs = np.array([[rel(f, r) for f, r in zip(_final(orc, name, dt0 / k, k * n0, mBC), ref)] for k in (1, 2, 4)])
    p = np.log2(errs[1] / errs[2])
    assert np.all(np.abs(p - ORDER[name]) < 0.5), (errs, p)


def test_rk_conduction_is_steady(orc):
    """No azimuthal structure, so no buoyancy torque: the velocities stay
    exactly zero.  The conductive profile (logarithmic in r) is steady up to
    its collocation residual, which the explicit diffusion integrates."""
    st = orc.seed_state(17, 8, ETA, 0, amp=0.0)
    run = orc.Run(17, 8, 1e5, 1.0, 1e-5, "RK4")
    run.set_state(*st)
    run.step(5)
    g = run.get_state()
    assert rel(g["tFR"], st[0]) < 1e-10
    for f in ("omgFR", "urFR", "upFR"):
        assert not g[f].any()
    d = run.diag()
    assert abs(d["Nu_t"] - 1) < 1e-8 and abs(d["Nu_b"] - 1) < 1e-8 and d["KE"] == 0.0


def _phys(f, Np):
    """Values on Np points of the real field with Fourier coefficients f
    (m = 0..Nm): f_0 + 2 Re sum_m f_m e^{2 pi i m j / Np}."""
    c = np.zeros(Np // 2 + 1, complex)
    c[:len(f)] = f
    c[0] = c[0].real
    return np.fft.irfft(c, Np) * Np


def test_rk_mean_flow_rhs_vs_numpy(orc):
    N, Nm = 17, 8
    run = orc.Run(N, Nm, 2e4, 1.0, 1e-5, "RK4")
    run.set_state(*orc.seed_state(N, Nm, ETA, 5, amp=0.5))
    run.step(2)                        # a moving state with upFC = chebtransform(u_phi)
    g = run.get_state()
    got = run.rk_mean_rhs()
    r = orc.grid(N, ETA)[0]
    x = cheb_points(N)
    Np = 3 * Nm

    def ddr(vals, k=1):                # d^k/dr^k of the interpolant, dr = dx / 2
        return npc.chebval(x, npc.chebder(npc.chebfit(x, vals, N - 1), k)) * 2.0 ** k

    u0 = g["upFR"][0].real
    udiff = ddr(u0, 2) - u0 / r ** 2 + ddr(u0) / r
    dup = np.array([ddr(g["upFR"][m].real) + 1j * ddr(g["upFR"][m].imag) for m in range(Nm + 1)])
    want = np.empty(N)
    for i in range(N):
        ur = _phys(g["urFR"][:, i], Np)
        up = _phys(g["upFR"][:, i], Np)
        du = _phys(dup[:, i], Np)
        want[i] = udiff[i] - (np.mean(ur * du) + np.mean(ur * up) / r[i])
    assert np.abs(got.imag).max() == 0.0
    assert np.abs(got.real - want).max() < 1e-9 * np.abs(want).max()
    # the nonlinear part is not negligible in this state
    assert np.abs(want - udiff).max() > 1e-3 * np.abs(want).max()


def test_rk_registry_and_limits(orc):
    with pytest.raises(ValueError):
        orc.Run(17, 8, 1e4, 1.0, 1e-5, ("RK", "NONE", "RK3"))
    run = orc.Run(17, 8, 1e4, 1.0, 1e-5, "RK4")
    run.set_state(*orc.seed_state(17, 8, ETA, 5))
    run.step(2)
    h = run.history(4, 4)              # store_checkpoint's (n_order, Nm+1, Nr) records
    assert h["imp_temp"].shape == (4, 9, 17) and not any(a.any() for a in h.values())
