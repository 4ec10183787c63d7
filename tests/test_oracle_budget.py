"""The oracle's diagnostics beyond KE and Nusselt: writeKE_spectral's
mean-flow energy, power budget and mean-flow balance (output.F90:122-324)
and calculate_spectra (output.F90:327-355).

Pinned by a numpy restatement from the fields alone: azimuthal means with
numpy's inverse real FFT on the Np = 3 Nm grid, radial derivatives and
integrals of numpy's Chebyshev interpolant (int f r dr, exact for it, which
is what Clenshaw-Curtis radInt computes).  The hidden inputs follow the
reference: the buoyancy power multiplies the current u_r by the temperature
of the last nonlinear pass (the module array tt; for IMEX, the state the
step started from), and du/dt of the mean flow is taken against upFR_prev
(IMEX: the start of the step; IMEXRK: the initial state, never updated on
that path).
"""
from __future__ import annotations

import numpy as np
import numpy.polynomial.chebyshev as npc
import pytest

from test_oracle_transforms import ETA, cheb_points

N, NM = 17, 8
FIELDS = ("tFR", "omgFR", "urFR", "upFR")


def _phys(f, Np):
    c = np.zeros(Np // 2 + 1, complex)
    c[:len(f)] = f
    c[0] = c[0].real
    return np.fft.irfft(c, Np) * Np


def _radint(f, r):
    x = cheb_points(len(r))
    c = npc.chebint(npc.chebfit(x, f * r, len(r) - 1))
    return 0.5 * (npc.chebval(1.0, c) - npc.chebval(-1.0, c))


def _ddr(vals, k=1):
    x = cheb_points(len(vals))
    fit = lambda v: npc.chebval(x, npc.chebder(npc.chebfit(x, v, len(v) - 1), k)) * 2.0 ** k
    return fit(vals.real) + 1j * fit(vals.imag)


def _numpy_budget(cur, t_pass, up_prev, dt, Ra, Pr, mBC, r):
    Np = 3 * NM
    pi = np.pi
    upavg = np.array([_phys(cur["upFR"][:, j], Np).mean() for j in range(N)])
    urt = np.array([np.mean(_phys(cur["urFR"][:, j], Np) * _phys(t_pass[:, j], Np)) for j in range(N)])
    w2 = np.array([np.mean(_phys(cur["omgFR"][:, j], Np) ** 2) for j in range(N)])
    up2 = (np.abs(cur["upFR"]) ** 2).sum(0)
    vis = 2 * pi * _radint(w2, r)
    if mBC == "SF":
        vis -= 4 * pi * 2 * (up2[-1] - up2[0])
    out = {"KE_mean": 2 * pi * _radint(upavg ** 2, r), "vis": vis, "buo": 2 * pi * Ra / Pr * _radint(urt, r)}
    u0 = cur["upFR"][0]
    dup = np.array([_ddr(cur["upFR"][m]) for m in range(NM + 1)])
    adv = np.array([np.mean(_phys(cur["urFR"][:, j], Np) * (_phys(dup[:, j], Np) + _phys(cur["upFR"][:, j], Np) / r[j]))
                    for j in range(N)])
    udiff = (_ddr(u0, 2) - u0 / r ** 2 + _ddr(u0) / r).real
    ub = np.stack([((u0 - up_prev) / dt).real, -adv, udiff], axis=1)
    return out, ub


@pytest.fixture(scope="module")
def developed(orc):
    """A flow with a mean flow of the order of the other modes."""
    A = orc.Run(N, NM, 1e6, 1.0, 1e-5, "CNAB2")
    A.set_state(*orc.seed_state(N, NM, ETA, 3, amp=1e-1))
    A.step(200)
    g = A.get_state()
    assert np.abs(g["upFR"][0]).max() > 0.1 * np.abs(g["upFR"]).max()
    return tuple(g[f] for f in FIELDS)


@pytest.mark.parametrize("scheme,mBC", [("CNAB2", "NS"), ("SBDF3", "SF"), ("ARS443", "NS"), ("BHR553", "SF"),
                                        ("RK4", "NS")])
def test_budget_vs_numpy(orc, developed, scheme, mBC):
    Ra, Pr = 1e6, 1.0
    dt = 1e-7 if scheme == "RK4" else 1e-6
    run = orc.Run(N, NM, Ra, Pr, dt, scheme, mBC=mBC)
    init = developed
    run.set_state(*init)
    run.step(3)
    before = run.get_state()
    run.step(1)
    cur = run.get_state()
    got, ub = run.budget()
    r = orc.grid(N, ETA)[0]
    imex = orc.SCHEMES[scheme][0] == "IMEX"
    # IMEX: the pass at the step's start saw the state before the step
    t_pass = before["tFR"] if imex else cur["tFR"]
    up_prev = before["upFR"][0] if imex else init[3][0]
    want, ubw = _numpy_budget(cur, t_pass, up_prev, dt, Ra, Pr, mBC, r)
    assert abs(got["t"] - run.diag()["t"]) == 0.0
    for k in ("KE_mean", "vis"):
        assert abs(got[k] - want[k]) <= 1e-10 * abs(want[k]), (k, got[k], want[k])
    if imex:
        assert abs(got["buo"] - want["buo"]) <= 1e-10 * abs(want["buo"])
    else:
        # the last pass saw a stage value inside the step, close to the new
        # state (the pass's temperature is pinned where it is a state: IMEX)
        assert abs(got["buo"] - want["buo"]) / abs(want["buo"]) < 1e-3
    for c in range(3):
        scale = np.abs(ubw[:, c]).max()
        assert np.abs(ub[:, c] - ubw[:, c]).max() <= 1e-9 * scale, (c, scheme)
    assert want["KE_mean"] > 0 and want["vis"] != 0 and np.abs(ubw).max() > 0


def test_budget_imex_uses_state_of_the_pass(orc, developed):
    """With the current temperature in place of the pass's, the buoyancy
    power misses: the quirk carries weight."""
    run = orc.Run(N, NM, 1e6, 1.0, 1e-6, "CNAB2")
    run.set_state(*developed)
    run.step(3)
    cur = run.get_state()
    got, _ = run.budget()
    r = orc.grid(N, ETA)[0]
    want, _ = _numpy_budget(cur, cur["tFR"], cur["upFR"][0], 1e-6, 1e6, 1.0, "NS", r)
    assert abs(got["buo"] - want["buo"]) > 1e-8 * abs(want["buo"])


def test_spectra_vs_numpy(orc):
    run = orc.Run(N, NM, 1e5, 1.0, 1e-5, "SBDF2")
    run.set_state(*orc.seed_state(N, NM, ETA, 8, amp=0.3))
    run.step(4)
    g = run.get_state()
    k, t = run.spectra()
    r = orc.grid(N, ETA)[0]
    for m in range(NM + 1):
        ek = np.abs(g["urFR"][m] ** 2 + g["upFR"][m] ** 2)
        assert abs(k[m] - 4 * np.pi * _radint(ek, r)) <= 1e-10 * max(k.max(), 1e-300)
        assert abs(t[m] - 4 * np.pi * _radint(np.abs(g["tFR"][m]), r)) <= 1e-10 * t.max()
    # conduction: only the mean temperature, no kinetic energy
    c = orc.Run(N, NM, 1e5, 1.0, 1e-5, "SBDF2")
    c.set_state(*orc.seed_state(N, NM, ETA, 0, amp=0.0))
    kc, tc = c.spectra()
    assert not kc.any() and tc[0] > 0 and not tc[1:].any()
