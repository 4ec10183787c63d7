"""Pins of the coupled right-hand side (buoyancy and flux-form advection) by
what the equations fix, independently of the oracle's code.

The reference advances, per Fourier mode m (rhs_create.F90:99-100, 214-254;
rhs_create_imexrk.F90:49-50, 224-225; fourierloop_imexrk.F90:280, 298-300,
484; rhs_create_rk.F90:57-60, 112-115)

    dT/dt = (1/Pr) lap_m T - [ i m/r (u_phi T) + d/dr (u_r T) + (u_r T)/r ]
    dw/dt =        lap_m w - [ i m/r (u_phi w) + d/dr (u_r w) + (u_r w)/r ]
                           - (Ra/Pr) i m T / r
    lap_m psi = -w,  u_r = i m psi / r,  u_phi = -d psi/dr,

with T = 1, 0 and psi = 0 at the inner and outer walls and d psi/dr = 0 (no
slip) or d2 psi/dr2 - (1/r) d psi/dr = 0 (stress free).  Three consequences of
these equations, computed here with numpy from the fields alone, must hold
for every time scheme to its temporal accuracy:

(a) linear growth rate: a small perturbation about the conductive state
    T0(r) = ln(r/rmax)/ln(rmin/rmax) along the leading eigenvector of the
    linearised operator (a generalised eigenproblem built from numpy
    Chebyshev collocation, the advection linearised as u_r dT0/dr) grows at
    that eigenvalue's rate;
(b) kinetic energy: d/dt 1/2 int |u|^2 dA = (Ra/Pr) int u_r T dA - int w^2 dA
    with no-slip walls (advection and pressure do no work);
(c) mean heat: the azimuthal mean of T changes only through the diffusive
    fluxes at the walls, d/dt int <T> r dr = (1/Pr) [r d<T>/dr] between the
    walls (the advective flux r <u_r T> vanishes there), written for the
    collocation rows the walls replace by their Dirichlet values.

Each check is shown to fail on a mistake planted in the oracle (a wrong
buoyancy sign, Ra in place of Ra/Pr, a wrong sign on (u_r f)/r): a slip of
that kind cannot pass these tests.
"""
from __future__ import annotations

import numpy as np
import numpy.polynomial.chebyshev as npc
import pytest
import scipy.linalg as sla

ETA = 0.35
RMIN = ETA / (1 - ETA)
RMAX = RMIN + 1.0


# ------------------------------------------------------------ numpy operators
def cheb_ops(N):
    """Gauss-Lobatto points in r (ascending, r = rmin + (1 + x)/2) and the
    collocation matrices of d/dr and d2/dr2 from numpy's Chebyshev series."""
    x = np.cos(np.pi * np.arange(N - 1, -1, -1) / (N - 1))
    Vi = np.linalg.inv(npc.chebvander(x, N - 1))
    d1 = np.stack([npc.chebval(x, npc.chebder(np.eye(N)[k])) for k in range(N)], 1) @ Vi
    d2 = np.stack([npc.chebval(x, npc.chebder(np.eye(N)[k], 2)) for k in range(N)], 1) @ Vi
    return RMIN + 0.5 * (1 + x), 2.0 * d1, 4.0 * d2


def cc_weights(N):
    """Clenshaw-Curtis weights of int_{-1}^{1} f dx on the same points, from
    the moments of T_k (int T_k = 2/(1-k^2) for even k, 0 for odd k)."""
    x = np.cos(np.pi * np.arange(N - 1, -1, -1) / (N - 1))
    mom = np.array([2.0 / (1 - k * k) if k % 2 == 0 else 0.0 for k in range(N)])
    return np.linalg.solve(npc.chebvander(x, N - 1).T, mom)


def linear_mode(N, m, Ra, Pr, mBC="NS"):
    """Leading eigenpair of the linearised equations for azimuthal mode m:
    B d/dt [T, w, psi] = A [T, w, psi] with the wall rows algebraic."""
    r, D1, D2 = cheb_ops(N)
    I = np.eye(N)
    L = D2 + np.diag(1 / r) @ D1 - m * m * np.diag(1 / r ** 2)
    dT0 = 1.0 / (r * np.log(RMIN / RMAX))
    A = np.zeros((3 * N, 3 * N), complex)
    B = np.zeros((3 * N, 3 * N))
    t, w, p = slice(0, N), slice(N, 2 * N), slice(2 * N, 3 * N)
    A[t, t] = L / Pr
    A[t, p] = -np.diag(1j * m / r * dT0)              # -u_r dT0/dr, u_r = i m psi / r
    B[t, t] = I
    A[w, w] = L
    A[w, t] = -(Ra / Pr) * np.diag(1j * m / r)        # buoyancy
    B[w, w] = I
    A[p, p] = L                                       # lap psi + w = 0
    A[p, w] = I
    for k in (0, N - 1):
        A[k], B[k] = 0, 0
        A[k, k] = 1.0                                 # T = 0 (perturbation)
        A[N + k], B[N + k] = 0, 0
        A[N + k, p] = D1[k] if mBC == "NS" else D2[k] - D1[k] / r[k]
        A[2 * N + k] = 0
        A[2 * N + k, 2 * N + k] = 1.0                 # psi = 0
    lam, V = sla.eig(A, B)
    ok = np.isfinite(lam)
    lam, V = lam[ok], V[:, ok]
    k = np.argmax(lam.real)
    v = V[:, k]
    return lam[k], v[t], v[w], v[p], r, D1


def parseval(f, g):
    """(1/2pi) int f g dphi of two real fields given by their modes 0..Nm
    (x = f_0 + 2 Re sum_{m>=1} f_m e^{i m phi}), per radial point."""
    return (f[0] * np.conj(g[0])).real + 2.0 * (f[1:] * np.conj(g[1:])).real.sum(0)


# ------------------------------------------------------------ drivers
def eigen_state(mod, N, Nm, m, lam_data, eps=1e-7):
    """Conduction plus eps times the eigenvector on mode m, as the four
    fields (tFR, omgFR, urFR, upFR)."""
    lam, th, om, ps, r, D1 = lam_data
    sc = eps / np.abs(th).max()
    th, om, ps = th * sc, om * sc, ps * sc
    t, o, u, w = (a.copy() for a in mod.seed_state(N, Nm, ETA, 0, amp=0.0))
    t[m], o[m], u[m], w[m] = th, om, 1j * m / r * ps, -(D1 @ ps)
    return (t, o, u, w), th


def growth_rate(make_run, state, th, m, dt, nsteps):
    """Growth rate of the projection of T_m on the eigenvector over the last
    step of nsteps (the run starts as a restart, so that the first step sees
    the explicit terms of its initial state)."""
    run = make_run(dt)
    run.set_state(*state)
    prev = None
    for _ in range(nsteps):
        run.step(1)
        c = np.vdot(th, run.get_state()["tFR"][m]) / np.vdot(th, th)
        g = None if prev is None else np.log(c / prev) / dt
        prev = c
    return g


def smooth_state(mod, N, Nm, amp=0.2, mmax=3):
    """Conduction plus smooth temperature perturbations on modes 1..mmax."""
    r = cheb_ops(N)[0]
    t, o, u, w = (a.copy() for a in mod.seed_state(N, Nm, ETA, 0, amp=0.0))
    for m in range(1, mmax + 1):
        t[m] = (amp / m) * np.exp(1j * m) * np.sin(np.pi * (r - RMIN)) * (1 + 0.5 * (r - RMIN))
    return t, o, u, w


def budgets(run, N, Ra, Pr, dt, nsteps=5):
    """Per step: (dKE/dt by central differences, buoyancy power, dissipation,
    d/dt of the interior heat content, the wall fluxes) in numpy from the
    run's fields."""
    r, D1, D2 = cheb_ops(N)
    w = cc_weights(N)
    L0 = D2 + np.diag(1 / r) @ D1
    area = lambda f: np.pi * np.sum(w * f * r)        # int f dA = 2 pi int f r dr, dr = dx/2
    rows = []
    for _ in range(nsteps):
        s = run.get_state()
        ke = 0.5 * area(parseval(s["urFR"], s["urFR"]) + parseval(s["upFR"], s["upFR"]))
        buo = (Ra / Pr) * area(parseval(s["urFR"], s["tFR"]))
        dis = area(parseval(s["omgFR"], s["omgFR"]))
        T0 = s["tFR"][0].real
        heat = 0.5 * np.sum((w * r * T0)[1:-1])       # interior rows: the walls keep T fixed
        f = parseval(s["urFR"], s["tFR"])             # <u_r T>, zero at the walls
        LT, Df = L0 @ T0, D1 @ f
        # int (r T0')' dr = [r T0'] and int (r f)' dr = 0 over all rows; the
        # wall rows' share of both integrals is taken off
        wall = 0.5 * sum(w[k] * r[k] * (LT[k] / Pr - Df[k]) for k in (0, N - 1))
        flux = (RMAX * (D1 @ T0)[-1] - RMIN * (D1 @ T0)[0]) / Pr - wall
        rows.append((ke, buo, dis, heat, flux))
        run.step(1)
    ke, buo, dis, heat, flux = np.array(rows).T
    c = slice(1, -1)
    return ((ke[2:] - ke[:-2]) / (2 * dt), buo[c], dis[c], (heat[2:] - heat[:-2]) / (2 * dt), flux[c])


def oracle_factory(orc, N, Nm, Ra, Pr, scheme, mBC="NS"):
    return lambda dt: orc.Run(N, Nm, Ra, Pr, dt, scheme, mBC=mBC, restart=True)


# ------------------------------------------------------------ (a) growth rate
# (scheme, dt, steps, relative tolerance).  CB3e's third explicit column sums
# to zero, so its explicit terms at that stage come from the previous stage's
# products (steptime_imexrk.F90:153-154, reproduced): first order there.
GROWTH = [("CNAB2", 1e-4, 40, 1e-5), ("CNAB3", 1e-4, 40, 1e-5), ("SBDF2", 1e-4, 40, 1e-5),
          ("SBDF3", 1e-4, 40, 1e-5), ("SBDF4", 1e-4, 40, 1e-5), ("ARS443", 1e-4, 40, 1e-5),
          ("BPR353", 1e-4, 40, 1e-5), ("ARS222", 1e-4, 40, 1e-5), ("BHR553", 1e-4, 40, 1e-5),
          ("KC343", 1e-4, 40, 1e-5), ("SMR333", 1e-4, 40, 1e-5), ("CB3e", 1e-4, 40, 2e-3),
          ("RK2", 2e-6, 400, 1e-5), ("RK4", 2e-6, 400, 1e-5)]
LIN = dict(N=33, Nm=4, m=2, Ra=1e5, Pr=7.0)


@pytest.fixture(scope="module")
def lin_mode():
    return {mBC: linear_mode(LIN["N"], LIN["m"], LIN["Ra"], LIN["Pr"], mBC) for mBC in ("NS", "SF")}


def test_linear_operator_sanity(lin_mode):
    """The eigenproblem itself: at Ra = 0 the leading temperature mode decays
    at the rate of (1/Pr) lap_m with Dirichlet walls; with buoyancy the
    leading mode at Ra = 1e5, Pr = 7 grows, and grows faster with larger Ra."""
    N, m, Pr = LIN["N"], LIN["m"], LIN["Pr"]
    r, D1, D2 = cheb_ops(N)
    L = (D2 + np.diag(1 / r) @ D1 - m * m * np.diag(1 / r ** 2))[1:-1, 1:-1]
    lam0 = linear_mode(N, m, 0.0, Pr)[0]
    lamT = np.linalg.eigvals(L / Pr)
    assert abs(lam0 - lamT[np.argmax(lamT.real)]) < 1e-9 * abs(lam0)
    lam = lin_mode["NS"][0]
    assert lam.real > 0 and abs(lam.imag) < 1e-8 * abs(lam)
    assert linear_mode(N, m, 2e5, Pr)[0].real > lam.real


@pytest.mark.parametrize("scheme,dt,steps,tol", GROWTH)
def test_linear_growth_rate(orc, lin_mode, scheme, dt, steps, tol):
    N, Nm, m, Ra, Pr = LIN["N"], LIN["Nm"], LIN["m"], LIN["Ra"], LIN["Pr"]
    lam = lin_mode["NS"][0]
    state, th = eigen_state(orc, N, Nm, m, lin_mode["NS"])
    g = growth_rate(oracle_factory(orc, N, Nm, Ra, Pr, scheme), state, th, m, dt, steps)
    assert abs(g - lam) < tol * abs(lam), (scheme, g, lam)


@pytest.mark.parametrize("scheme", ["CNAB2", "SBDF3", "ARS443", "BHR553"])
def test_linear_growth_rate_stress_free(orc, lin_mode, scheme):
    N, Nm, m, Ra, Pr = LIN["N"], LIN["Nm"], LIN["m"], LIN["Ra"], LIN["Pr"]
    lam = lin_mode["SF"][0]
    assert abs(lam - lin_mode["NS"][0]) > 0.1 * abs(lam)     # the wall condition matters
    state, th = eigen_state(orc, N, Nm, m, lin_mode["SF"])
    g = growth_rate(oracle_factory(orc, N, Nm, Ra, Pr, scheme, "SF"), state, th, m, 1e-4, 40)
    assert abs(g - lam) < 1e-5 * abs(lam), (scheme, g, lam)


@pytest.mark.parametrize("mutation", [1, 2, 3])
@pytest.mark.parametrize("scheme", ["CNAB2", "SBDF3", "ARS443", "BHR553", "RK4"])
def test_linear_growth_rate_catches_planted_mistakes(orc, lin_mode, scheme, mutation):
    N, Nm, m, Ra, Pr = LIN["N"], LIN["Nm"], LIN["m"], LIN["Ra"], LIN["Pr"]
    dt, steps = (2e-6, 400) if scheme == "RK4" else (1e-4, 40)
    lam = lin_mode["NS"][0]
    state, th = eigen_state(orc, N, Nm, m, lin_mode["NS"])
    with orc.mutation(mutation):
        g = growth_rate(oracle_factory(orc, N, Nm, Ra, Pr, scheme), state, th, m, dt, steps)
    assert abs(g - lam) > 0.1 * abs(lam), (scheme, mutation, g, lam)


# ------------------------------------------------------------ (b), (c) budgets
BUD = dict(N=33, Nm=16, Ra=1e5, Pr=0.7)
BUDGET_SCHEMES = [("CNAB2", 2e-5, 200), ("SBDF3", 2e-5, 200), ("ARS443", 2e-5, 200), ("BPR353", 2e-5, 200),
                  ("BHR553", 2e-5, 200), ("SMR333", 2e-5, 200), ("RK4", 2e-6, 2000)]


def oracle_budgets(orc, scheme, dt, warm, mutation=0):
    N, Nm, Ra, Pr = BUD["N"], BUD["Nm"], BUD["Ra"], BUD["Pr"]
    with orc.mutation(mutation):
        run = orc.Run(N, Nm, Ra, Pr, dt, scheme, restart=True)
        run.set_state(*smooth_state(orc, N, Nm))
        run.step(warm)
        return budgets(run, N, Ra, Pr, dt)


def energy_err(dke, buo, dis):
    return np.abs(dke - (buo - dis)).max() / max(np.abs(buo).max(), np.abs(dis).max())


def heat_err(dh, flux):
    return np.abs(dh - flux).max() / np.abs(flux).max()


@pytest.mark.parametrize("scheme,dt,warm", BUDGET_SCHEMES)
def test_energy_and_heat_budgets(orc, scheme, dt, warm):
    dke, buo, dis, dh, flux = oracle_budgets(orc, scheme, dt, warm)
    assert buo.min() > 0.2 * dis.max()          # both terms carry weight in this flow
    assert energy_err(dke, buo, dis) < 1e-3, (scheme, dke, buo - dis)
    assert heat_err(dh, flux) < 1e-3, (scheme, dh, flux)


@pytest.mark.parametrize("scheme,dt,warm", [BUDGET_SCHEMES[i] for i in (0, 1, 2, 4, 6)])
def test_budgets_catch_planted_mistakes(orc, scheme, dt, warm):
    """Energy catches every planted slip (the advection slip does work on
    the flow); the heat balance catches the advection slip."""
    for mutation in (1, 2, 3):
        dke, buo, dis, dh, flux = oracle_budgets(orc, scheme, dt, warm, mutation)
        assert energy_err(dke, buo, dis) > 0.03, (scheme, mutation)
        if mutation == 3:
            assert heat_err(dh, flux) > 0.5, (scheme, mutation)
