"""Variable dt (l_vartimestep) in the oracle: compute_new_dt
(steptime.F90:313-370, steptime_imexrk.F90:248-308), get_dtvals
(nonlin.F90:93-106) and the re-factorisation when dt changes
(fourierloop.F90:86-94).

Pinned by what the mathematics fixes:
* At rest (Ra = 0) the CFL limit is infinite, so dt goes from the first
  step's dt to dt_max.  On a diffusion eigenmode, CN then multiplies by
  (1 + z/2)/(1 - z/2) with z = dt_k lambda of each step's own dt, and the
  BDF-k schemes follow the variable-step BDF recurrence whose coefficients
  are the derivative at t_{n+1} of the Lagrange interpolant through the last
  k + 1 time levels (computed here from the nodes): a wrong dt_array order,
  wrong ratios or operators built at the wrong steps miss both.
* With a moving fluid, each step's dt is CFL times the smallest of
  dr/max|u_r| and (2/3) pi r/(Nm max|u_phi|) over the levels of the state the
  step starts from, the velocity maxima taken here from numpy's inverse real
  FFT of that state.
"""
from __future__ import annotations

import numpy as np
import pytest

from conftest import rel
from test_oracle_schemes import ETA, M_LIN, N_LIN, NM_LIN, eigen_start


def _run_rest(orc, scheme, dt, dt_max, steps, restart):
    state, v, lam = eigen_start(N_LIN, M_LIN, NM_LIN, orc)
    run = orc.Run(N_LIN, NM_LIN, 0.0, 1.0, dt, scheme, restart=restart, vardt=True, CFL=0.5, dt_coef=1.0,
                  dt_max=dt_max)
    run.set_state(*state)
    T, dts = [], []
    for _ in range(steps):
        run.step(1)
        T.append(run.get_state()["tFR"][M_LIN])
        dts.append(run.dt())
    return T, dts, v, lam


def test_vardt_cn_eigenmode(orc):
    dt, dt_max = 1e-3, 2.5e-3
    T, dts, v, lam = _run_rest(orc, "CNAB2", dt, dt_max, 5, restart=False)
    # fresh start: the first step keeps dt (steptime.F90:224-231), then the
    # infinite CFL limit is capped at dt_max
    assert dts == [dt] + [dt_max] * 4
    y = v
    for n in range(5):
        z = dts[n] * lam
        y = y * (1 + z / 2) / (1 - z / 2)
        assert rel(T[n], y) < 1e-10, n


def _lagrange_d0(times):
    """Weights c_j with p'(t_0) = sum_j c_j p(t_j) for the interpolant p
    through the nodes times[0..k]."""
    t = np.asarray(times, float)
    k = len(t)
    c = np.zeros(k)
    for j in range(k):
        others = [t[i] for i in range(k) if i != j]
        den = np.prod([t[j] - o for o in others])
        s = 0.0
        for a in range(len(others)):
            s += np.prod([t[0] - o for b, o in enumerate(others) if b != a])
        c[j] = s / den
    return c


@pytest.mark.parametrize("scheme,k", [("SBDF2", 2), ("SBDF3", 3), ("SBDF4", 4)])
def test_vardt_bdf_eigenmode(orc, scheme, k):
    """Variable-step BDF-k: y_{n+1} = sum_j (-c_j/c_0) y_{n+1-j} / (1 - lam/c_0')
    with c the Lagrange derivative weights of the step's time levels.  c_0'
    is c_0 of the step that last built the operator: mat_build runs only on
    the first step and when dt_array(1) /= dt_array(2) (fourierloop.F90:86-87),
    so after a dt change the operator keeps that step's wt_lhs * dt (= 1/c_0)
    while the right-hand-side weights follow the new ratios -- a quirk of the
    reference, kept."""
    dt, dt_max, steps = 1e-3, 2.5e-3, 6
    T, dts, v, lam = _run_rest(orc, scheme, dt, dt_max, steps, restart=True)
    # restart: the first step already takes the CFL-limited dt; dt_array(2..)
    # keep the first dt (steptime.F90:349-356)
    assert dts == [dt_max] * steps
    dta = [dt] * 8                       # the dt_array history, newest first
    hist = [v] * k                       # identical checkpoints of the state
    c0_op = None
    for n in range(steps):
        # first step of a restart: dt_array(1) only; later: shift every slot
        dta = [dts[0]] + dta[1:] if n == 0 else [dts[n]] + dta[:7]
        times = np.concatenate([[0.0], -np.cumsum(dta[:k])])
        c = _lagrange_d0(times)
        if n == 0 or dta[0] != dta[1]:
            c0_op = c[0]
        y = sum(-c[j] / c[0] * hist[j - 1] for j in range(1, k + 1)) / (1 - lam / c0_op)
        assert rel(T[n], y) < 1e-10, (scheme, n)
        hist = [y] + hist[:-1]


def _phys_max(f, Np):
    """max |x| of the real field with Fourier coefficients f (m = 0..Nm) on
    Np points: x_j = f_0 + 2 Re sum_m f_m e^{2 pi i m j / Np}."""
    c = np.zeros(Np // 2 + 1, complex)
    c[:len(f)] = f
    c[0] = c[0].real
    return np.abs(np.fft.irfft(c, Np) * Np).max()


def test_vardt_cfl_limit_from_velocities(orc):
    Nr, Nm, CFL, dt_max = 17, 8, 0.5, 1e-2
    A = orc.Run(Nr, Nm, 1e6, 1.0, 1e-5, "CNAB2", eta=ETA)
    A.set_state(*orc.seed_state(Nr, Nm, ETA, 3, amp=1e-1))
    A.step(300)                                          # a moving fluid
    g = A.get_state()
    B = orc.Run(Nr, Nm, 1e6, 1.0, 1e-5, "CNAB2", eta=ETA, vardt=True, CFL=CFL, dt_coef=1.0, dt_max=dt_max)
    B.set_state(g["tFR"], g["omgFR"], g["urFR"], g["upFR"])
    r = orc.grid(Nr, ETA)[0]
    dr = np.r_[r[1] - r[0], np.diff(r)]
    Np = 3 * Nm
    B.step(1)
    assert B.dt() == 1e-5
    for _ in range(6):
        s = B.get_state()                                # the state the next step starts from
        dtv = [min(dr[i] / _phys_max(s["urFR"][:, i], Np),
                   (2 / 3) * r[i] * np.pi / (Nm * _phys_max(s["upFR"][:, i], Np))) for i in range(Nr)]
        B.step(1)
        want = min(CFL * min(dtv), dt_max)
        assert abs(B.dt() - want) <= 1e-13 * want
    assert B.dt() < dt_max                               # the CFL limit, not the cap, decided


def test_vardt_capped_run_equals_constant_run(orc):
    """dt = dt_max below the CFL limit: every step keeps dt_max, and the run
    is the constant-dt run, bit for bit."""
    Nr, Nm, dt = 17, 8, 1e-6
    st = orc.seed_state(Nr, Nm, ETA, 4, amp=1e-2)
    for scheme in ("CNAB2", "SBDF3", "ARS443"):
        a = orc.Run(Nr, Nm, 1e5, 1.0, dt, scheme, eta=ETA, vardt=True, CFL=0.5, dt_coef=1.0, dt_max=dt)
        b = orc.Run(Nr, Nm, 1e5, 1.0, dt, scheme, eta=ETA)
        a.set_state(*st)
        b.set_state(*st)
        a.step(6)
        b.step(6)
        assert a.dt() == dt
        ga, gb = a.get_state(), b.get_state()
        for f in ("tFR", "omgFR", "urFR", "upFR"):
            assert np.array_equal(ga[f], gb[f]), (scheme, f)


def test_vardt_rk_dt_sequence(orc):
    """The RK path (steptime_rk.F90:181-236): dt_array is reset to the
    namelist dt at every step, the first step keeps dt, later steps take the
    optimize branches against dt with dt_cal including the diffusive limits
    dr^2 and ((2/3) pi r / Nm)^2 (nonlin.F90:102-104), recomputed here from
    the grid.  At rest the CFL limits are infinite, so the diffusive ones
    decide; each of the three branches is taken once."""
    N, Nm, CFL, dt_max = 17, 8, 0.5, 1e-2
    r = orc.grid(N, ETA)[0]
    dr = np.r_[r[1] - r[0], np.diff(r)]
    dt_cal = min(np.min(dr ** 2), np.min(((2 / 3) * (r * np.pi) / Nm) ** 2))
    for dt, dt_coef in ((1e-6, 1.0), (1e-3, 1.0), (1e-6, 1e-3)):
        run = orc.Run(N, Nm, 0.0, 1.0, dt, "RK4", vardt=True, CFL=CFL, dt_coef=dt_coef, dt_max=dt_max)
        st = orc.seed_state(N, Nm, ETA, 0, amp=0.0)
        run.set_state(*st)
        dts = []
        for _ in range(3):
            run.step(1)
            dts.append(run.dt())
        dt2 = min(0.5 * (1 / dt_coef + 1) * CFL * dt_cal, dt_max)
        if dt >= dt_cal:
            want = dt2
        elif dt_coef * dt < dt_cal:
            want = max(dt, dt2)
        else:
            want = dt                     # no branch: dt stays
        assert dts[0] == dt
        assert all(abs(d - want) <= 1e-15 * want for d in dts[1:]), (dt, dt_coef, dts, want)
