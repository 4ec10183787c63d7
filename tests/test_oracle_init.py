"""The reference's fresh-start initial state in the oracle: init_perturbation
(init.F90:277-399) with its Gaspari-Cohn pieces (init.F90:937-985), and the
dt_array a restart restores from its checkpoints (init.F90:464-612).

Pinned by what the definitions fix, not by restating them:
* the compact pieces of Gaspari & Cohn (1999, eq. 4.7): the edge piece
  vanishes at twice its half-width (so the bubble meets the wall values), it
  joins the middle piece at the half-width, and the middle piece tends to the
  centre value at zero distance (so the bubble's peak is ampT);
* n_init > 0: a cos(n phi) perturbation has the single Fourier coefficient
  (ampT/2)(rmax - r)(rmin - r) on mode n;
* n_init < 0: zero outside 1 < phi < 3 and the Gaussian's closed form at its
  centre;
* the spectral state is numpy's real FFT of the physical field, level by
  level, scaled by 1/Np and cut to modes 0..Nm.
"""
from __future__ import annotations

import numpy as np
import pytest

from conftest import rel

ETA = 0.35
RMIN = ETA / (1 - ETA)
RMAX = 1 / (1 - ETA)


def grid(N):
    return RMIN + 0.5 * (1 + np.cos(np.pi * np.arange(N - 1, -1, -1) / (N - 1)))


def tcon(r):
    return np.log(r / RMAX) / np.log(RMIN / RMAX)


@pytest.mark.parametrize("c,L", [(0.25, 0.035355), (0.2, 0.05), (0.31, 0.0354)])
def test_gaspari_pieces(orc, c, L):
    g = orc.gaspari
    scale = g("center", 0.0, c, L)
    assert scale > 0
    assert abs(g("edge", 2 * c, c, L)) < 1e-13 * scale               # compact support
    assert abs(g("edge", c, c, L) - g("middle", c, c, L)) < 1e-12 * scale   # the pieces join
    assert abs(g("middle", 1e-7 * c, c, L) / scale - 1) < 1e-5      # the centre value is the limit
    # decreasing away from the centre
    d = np.linspace(0.05 * c, 1.95 * c, 20)
    vals = [g("middle", x, c, L) if x < c else g("edge", x, c, L) for x in d]
    assert np.all(np.diff(vals) < 0)


@pytest.mark.parametrize("N,Nm,n", [(17, 8, 3), (33, 16, 16), (32, 32, 1)])
def test_single_mode_perturbation(orc, N, Nm, n):
    ampT = 0.01
    t, o, u, w = orc.init_perturbation(N, Nm, ETA, n, ampT)
    r = grid(N)
    want = np.zeros((Nm + 1, N), complex)
    want[0] = tcon(r)
    want[n] += 0.5 * ampT * (RMAX - r) * (RMIN - r)
    assert np.abs(t - want).max() < 1e-15
    assert not o.any() and not u.any() and not w.any()


def test_gaussian_perturbation(orc):
    N, Nm, ampT = 33, 32, 0.01
    Np = 3 * Nm
    tt = orc.perturbation_field(N, Nm, ETA, -1, ampT)
    r = grid(N)
    phi = np.cumsum(np.r_[0.0, np.full(Np - 1, 2 * np.pi / Np)])
    pert = tt - tcon(r)[None, :]
    outside = (phi <= 1.0) | (phi >= 3.0)
    assert np.abs(pert[outside]).max() < 1e-15
    # the centre (phi nearest 2, r = rmax - 1/2): -(rmax - r)(rmin - r) ampT / (2 pi sigma^2) exp(-d^2/sigma^2)
    i = np.argmin(np.abs(phi - 2.0))
    j = N // 2
    sig = 0.25
    d2 = (r[j] - (RMAX - 0.5)) ** 2 + (phi[i] - 2.0) ** 2
    want = -(RMAX - r[j]) * (RMIN - r[j]) * ampT / (2 * np.pi * sig * sig) * np.exp(-d2 / sig ** 2)
    assert abs(pert[i, j] - want) < 1e-14
    assert pert[i, j] > 0.9 * ampT / (2 * np.pi * sig * sig) * 0.25   # the bump is positive and of its size


@pytest.mark.parametrize("N,Nm", [(33, 16), (32, 32), (65, 64)])
def test_bubble_perturbation(orc, N, Nm):
    ampT = 0.01
    Np = 3 * Nm
    tt = orc.perturbation_field(N, Nm, ETA, 0, ampT)
    r = grid(N)
    pert = tt - tcon(r)[None, :]
    ir = np.argmin(np.abs(r - 0.5 * (RMIN + RMAX)))
    # peak ampT at the grid point nearest the mid radius and phi = pi
    assert abs(pert[Np // 2, ir] - ampT) < 1e-15
    assert np.unravel_index(np.argmax(pert), pert.shape) == (Np // 2, ir)
    # compact support: the walls keep the conductive values
    assert np.abs(pert[:, 0]).max() < 1e-16 and np.abs(pert[:, -1]).max() < 1e-16
    assert abs(tt[0, 0] - 1.0) < 1e-15 and abs(tt[0, -1]) < 1e-15
    # symmetric about phi = pi
    assert np.abs(pert[1:Np // 2] - pert[Np - 1:Np // 2:-1]).max() < 1e-15


@pytest.mark.parametrize("n_init", [0, -1, 5])
def test_spectral_state_is_forward_fft(orc, n_init):
    N, Nm, ampT = 33, 16, 0.01
    Np = 3 * Nm
    tt = orc.perturbation_field(N, Nm, ETA, n_init, ampT)
    t = orc.init_perturbation(N, Nm, ETA, n_init, ampT)[0]
    want = (np.fft.rfft(tt, axis=0) / Np)[:Nm + 1]
    assert rel(t, want) < 1e-14


# --------------------------------------------------------------- restart dt
@pytest.mark.parametrize("scheme,k", [("SBDF2", 2), ("SBDF3", 3), ("SBDF4", 4)])
def test_restart_dt_array_weights(orc, scheme, k):
    """At rest, a variable-dt BDF-k restart from k checkpoints of a diffusion
    eigenmode with dt_array (d_1, ..., d_k) restored: its first step puts
    the CFL dt (here dt_max) in entry 1 and keeps d_2..d_k
    (steptime.F90:344-348), so the step is the variable-step BDF-k of the
    time levels 0, -dt_max, -dt_max - d_2, ... (Lagrange derivative
    weights).  Without the restored entries the namelist dt stands in."""
    from test_oracle_schemes import M_LIN, N_LIN, NM_LIN, eigen_start
    from test_vardt import _lagrange_d0
    state, v, lam = eigen_start(N_LIN, M_LIN, NM_LIN, orc)
    dt, dt_max = 1e-3, 2.5e-3
    dta = [1.7e-3, 1.3e-3, 2.1e-3, 0.9e-3][:k]
    amps = [1.0, 0.95, 0.91, 0.88][:k]               # newest first
    states = []
    for a in reversed(amps):                         # oldest first
        st = [f.copy() for f in state]
        st[0][M_LIN] = a * v
        states.append(tuple(st))
    hist = [a * v for a in amps]

    def first_step(with_dt):
        run = orc.Run(N_LIN, NM_LIN, 0.0, 1.0, dt, scheme, restart=True, vardt=True, CFL=0.5, dt_coef=1.0,
                      dt_max=dt_max)
        run.restart(states, 0.0)
        if with_dt:
            run.restart_dt(dta)
        run.step(1)
        return run.get_state()["tFR"][M_LIN], run.dt()

    for with_dt in (True, False):
        T, d = first_step(with_dt)
        assert d == dt_max
        older = dta[1:] if with_dt else [dt] * (k - 1)
        times = np.concatenate([[0.0], -np.cumsum([dt_max] + older)])
        c = _lagrange_d0(times)
        y = sum(-c[j] / c[0] * hist[j - 1] for j in range(1, k + 1)) / (1 - lam / c[0])
        assert rel(T, y) < 1e-10, (scheme, with_dt)
    assert rel(first_step(True)[0], first_step(False)[0]) > 1e-6


def _vardt_states(orc, scheme, steps):
    """An uninterrupted variable-dt run: per step its state, time and dt."""
    run = orc.Run(17, 8, 2e4, 1.0, 1e-4, scheme, restart=True, vardt=True, CFL=0.5, dt_coef=1.0, dt_max=4e-4)
    run.set_state(*orc.seed_state(17, 8, ETA, 3, amp=0.5))
    out = []
    for _ in range(steps):
        run.step(1)
        s = run.get_state()
        out.append(((s["tFR"], s["omgFR"], s["urFR"], s["upFR"]), run.diag()["t"], run.dt()))
    return out


@pytest.mark.parametrize("scheme,K", [("CNAB2", 2), ("SBDF3", 3), ("SBDF4", 4)])
def test_restart_continues_variable_dt_run(orc, scheme, K):
    """Restarted from K checkpoints of an uninterrupted variable-dt run, at a
    step where that run rebuilds its operator (dt changes) and where the
    restart's dt_array (entry 1 replaced by the new dt) equals the run's own,
    the restart continues it; the namelist dt in the older entries does not."""
    steps = 12
    ref = _vardt_states(orc, scheme, steps)
    dts = [r[2] for r in ref]
    # first restarted step cut+1 changes dt, and the K dts before it agree
    cut = next(c for c in range(K, steps - 2) if dts[c] != dts[c - 1] and len(set(dts[c - K:c])) == 1)
    states = [ref[cut - K + j][0] for j in range(K)]
    dta = [ref[cut - 1 - k][2] for k in range(K)] + [1e-4] * (max(K, 3 if scheme == "CNAB2" else K) - K)

    def restarted(with_dt):
        run = orc.Run(17, 8, 2e4, 1.0, 1e-4, scheme, restart=True, vardt=True, CFL=0.5, dt_coef=1.0, dt_max=4e-4)
        run.restart(states, ref[cut - 1][1])
        if with_dt:
            run.restart_dt(dta)
        run.step(steps - cut)
        return run.get_state()["tFR"]

    want = ref[-1][0][0]
    assert rel(restarted(True), want) < 1e-12
    assert rel(restarted(False), want) > 1e-6
