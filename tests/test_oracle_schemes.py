"""Pins of the oracle's time stepping against textbook numerics.

* Tableaux of the stiffly accurate IMEXRK schemes (rhs_update_wts_imp/exp of
  timeschemes.F90): order conditions of each part to the published order,
  shared abscissae and the order-3 coupling conditions, stiffly accurate last
  rows.
* Whole steps with Ra = 0 and zero velocity, where only the implicit part acts
  on each temperature mode: one discrete eigenvector of the collocation
  operator (built here with numpy, not the oracle) must be multiplied by the
  scheme's amplification factor -- CN (1 + z/2)/(1 - z/2), the BDFk recurrence
  with textbook coefficients, and R(z) = 1 + z b^T (I - z A)^{-1} 1 of the
  IMEXRK tableau.  This pins mat_build, the RHS assembly of rhs_create.F90 and
  rhs_create_imexrk.F90, the Chebyshev transforms and the solves together.
* The fresh-start quirk of zero-initialised histories (timeschemes.F90:111-113):
  BDF3 scales the first step by 18/11.
* Pure conduction stays put with zero velocity and Nu = 1 (output.F90:232-265).
* Temporal order of convergence of the coupled nonlinear problem.
"""
from __future__ import annotations

import numpy as np
import numpy.polynomial.chebyshev as npc
import pytest

from conftest import rel
from test_oracle_transforms import ETA, cheb_points


# ------------------------------------------------------------------ tableaux
# the stiffly accurate IMEXRK schemes of the registry: (name, stages, order)
RK_SA = [("ARS222", 3, 2), ("PC2", 4, 2), ("SSP332", 4, 2), ("SSP543", 6, 3), ("ARS443", 5, 3),
         ("BPR353", 5, 3), ("FW53", 5, 3), ("CK222", 3, 2), ("BPR442", 5, 2)]


@pytest.mark.parametrize("name,S,order", RK_SA)
def test_imexrk_tableau_order_conditions(orc, name, S, order):
    """Order conditions of each part up to the published order, with shared
    abscissae (so the coupling conditions of order 2 hold) and the order-3
    coupling conditions: a mistyped coefficient breaks one of them."""
    info = orc.scheme_info("IMEXRK", name, "exp" + name)
    aD, aA, bD, bA = info["aD"], info["aA"], info["bD"], info["bA"]
    assert len(bD) == S and info["sa"]
    cD, cA = aD.sum(1), aA.sum(1)
    assert np.allclose(cD, cA, atol=1e-15)           # shared abscissae
    c = cA
    assert aD[0, 0] == 0.0
    assert np.allclose(np.triu(aA), 0.0)             # explicit part strictly lower
    assert np.allclose(np.triu(aD, 1), 0.0)          # implicit part lower
    assert np.array_equal(aD[-1], bD) and np.array_equal(aA[-1], bA)   # stiffly accurate
    for b, A in ((bD, aD), (bA, aA)):
        assert abs(b.sum() - 1) < 1e-14
        assert abs(b @ c - 0.5) < 1e-14
        if order >= 3:
            assert abs(b @ c ** 2 - 1 / 3) < 1e-14
            assert abs(b @ A @ c - 1 / 6) < 1e-14
    if order >= 3:
        assert abs(bD @ aA @ c - 1 / 6) < 1e-14
        assert abs(bA @ aD @ c - 1 / 6) < 1e-14
    else:
        # second order only: the third-order condition fails for at least one part
        assert max(abs(bD @ c ** 2 - 1 / 3), abs(bA @ c ** 2 - 1 / 3), abs(bD @ aD @ c - 1 / 6),
                   abs(bA @ aA @ c - 1 / 6)) > 1e-3


# the IMEXRK schemes that end with Assembly_stage: (name, stages, order of the
# conditions their tableaux meet with shared abscissae; 1 for IJ3 and PR433,
# whose implicit and explicit abscissae differ)
RK_ASSEMBLY = [("EUL111", 2, 1), ("ARS232", 3, 2), ("ARS233", 3, 3), ("ARS343", 4, 3), ("MARS343", 4, 3),
               ("IJ3", 4, 1), ("PR433", 5, 1), ("DBM453", 5, 3), ("KC343", 4, 3), ("KC564", 6, 3),
               ("KC774", 7, 3), ("KC885", 8, 3), ("CFN343", 4, 3), ("BHR553", 5, 3), ("CFN564", 6, 3)]


@pytest.mark.parametrize("name,S,order", RK_ASSEMBLY)
def test_imexrk_assembly_tableaux(orc, name, S, order):
    """Not stiffly accurate by the reference's flags, so the step ends with
    Assembly_stage; order conditions of each part and the coupling ones."""
    info = orc.scheme_info("IMEXRK", name, "exp" + name)
    aD, aA, bD, bA = info["aD"], info["aA"], info["bD"], info["bA"]
    assert len(bD) == S and not info["sa"]
    assert np.allclose(np.triu(aA), 0.0) and np.allclose(np.triu(aD, 1), 0.0) and aD[0, 0] == 0.0
    for b in (bD, bA):
        assert abs(b.sum() - 1) < 1e-13
    if order >= 2:
        c = aA.sum(1)
        assert np.allclose(aD.sum(1), c, atol=1e-13)
        for b, A in ((bD, aD), (bA, aA)):
            assert abs(b @ c - 0.5) < 1e-13
            if order >= 3:
                assert abs(b @ c ** 2 - 1 / 3) < 1e-13
                assert abs(b @ A @ c - 1 / 6) < 1e-13
        if order >= 3:
            assert abs(bD @ aA @ c - 1 / 6) < 1e-13
            assert abs(bA @ aD @ c - 1 / 6) < 1e-13


@pytest.mark.parametrize("name", [r[0] for r in RK_ASSEMBLY])
def test_imexrk_assembly_amplification(orc, name):
    """At rest the temperature stages are Y = 1 + z A' Y (A' with the aD(2,2)
    operator for every stage) and Assembly_stage gives y_{n+1} = 1 + z bD^T Y."""
    dt = 2e-3
    info = orc.scheme_info("IMEXRK", name, "exp" + name)
    S = len(info["bD"])
    A = info["aD"].copy()
    for k in range(1, S):
        A[k, k] = info["aD"][1, 1]
    state, v, lam = eigen_start(N_LIN, M_LIN, NM_LIN, orc)
    z = dt * lam
    R = 1 + z * info["bD"] @ np.linalg.solve(np.eye(S) - z * A, np.ones(S))
    T = advance(orc, N_LIN, NM_LIN, dt, name, state, 3)
    for n, Tn in enumerate(T, 1):
        assert rel(Tn[M_LIN], R ** n * v) < 1e-10


# IMEXRK schemes whose stages are solved with their own diagonal (diag_diff):
# (name, stages, stiffly accurate by the reference's flags)
RK_DIAG = [("SMR333", 4, True), ("CB3e", 4, False), ("CB3f", 4, False), ("CB4", 6, False), ("LZ3L1", 5, True),
           ("LZ3A4a", 5, True), ("LZ4A1", 7, True)]


@pytest.mark.parametrize("name,S,sa", RK_DIAG)
def test_imexrk_own_diagonal_amplification(orc, name, S, sa):
    """Stage k solved with the operator of its own aD(k,k)
    (fourierloop_imexrk.F90:134-158): at rest Y = 1 + z aD Y, and the step
    gives Y_S (stiffly accurate) or 1 + z bD^T Y (Assembly_stage); the
    tableaux meet the first- and second-order conditions with shared
    abscissae."""
    info = orc.scheme_info("IMEXRK", name, "exp" + name)
    aD, aA, bD, bA = info["aD"], info["aA"], info["bD"], info["bA"]
    assert len(bD) == S and info["sa"] == sa
    c = aA.sum(1)
    assert np.allclose(aD.sum(1), c, atol=1e-12)
    for b in (bD, bA):
        assert abs(b.sum() - 1) < 1e-13 and abs(b @ c - 0.5) < 1e-12
    dt = 2e-3
    state, v, lam = eigen_start(N_LIN, M_LIN, NM_LIN, orc)
    z = dt * lam
    Y = np.linalg.solve(np.eye(S) - z * aD, np.ones(S))
    R = Y[-1] if sa else 1 + z * bD @ Y
    T = advance(orc, N_LIN, NM_LIN, dt, name, state, 3)
    for n, Tn in enumerate(T, 1):
        assert rel(Tn[M_LIN], R ** n * v) < 1e-10


@pytest.mark.parametrize("mBC", ["NS", "SF"])
def test_zero_weight_vorticity_operator_is_singular(orc, mBC):
    """mat_build_IRK at aD(k,k) = 0 (LZ3A4a's stages 2 and 4): omega's
    interior rows reduce to omega(r_i) = rhs, and no row involves omega at the
    walls, so the matrix loses rank 2: omega = (a + b x) T'_{N-1}(x), zero at
    every interior Gauss-Lobatto point and free at the walls, with psi = 0,
    is in its null space (checked with a = b, the cardinal polynomial of the
    wall x = 1).  Why the library refuses the scheme."""
    N = 17
    for n in (1, 3):
        _, AF, _ = orc.mat_build(N, ETA, n, 0.0, 1.0, mBC, irk_order=True)
        s = np.linalg.svd(AF, compute_uv=False)
        assert s[-3] / s[0] > 1e-8 and s[-2] / s[0] < 1e-14
        _, AF1, _ = orc.mat_build(N, ETA, n, 1e-3, 1.0, mBC, irk_order=True)
        s1 = np.linalg.svd(AF1, compute_uv=False)
        assert s1[-1] / s1[0] > 1e-12
        # omega = the wall bump, psi = 0: AF [0; c] = 0 with c the Chebyshev
        # coefficients (end columns halved and scaled like the operator's)
        x = cheb_points(N)
        bump = np.where(np.abs(np.abs(x) - 1) < 1e-14, 1.0, 0.0) * (x > 0)
        c = np.polynomial.chebyshev.chebfit(x, bump, N - 1)
        c[0] *= 2.0
        c[-1] *= 2.0
        v = np.concatenate([np.zeros(N), c / np.sqrt(2.0 / (N - 1))])
        assert np.abs(AF @ v).max() < 1e-10 * np.abs(AF).max() * np.abs(v).max()


def test_scheme_registry(orc):
    assert orc.scheme_info("IMEX", "CN", "AB2")["n_imp"] >= 1
    assert orc.scheme_info("IMEX", "BDF3", "AB3")["n_exp"] == 3
    with pytest.raises(ValueError):
        orc.scheme_info("IMEX", "BDF7", "AB2")
    with pytest.raises(ValueError):
        orc.scheme_info("IMEXRK", "ARS443", "expBPR353")


# --------------------------------------------------- linear amplification
def collocation_L(N, m):
    """Interior block of D2 + D/r - m^2/r^2 on the Gauss-Lobatto grid
    (Dirichlet walls), from numpy's Chebyshev Vandermonde matrices."""
    x = cheb_points(N)
    V = npc.chebvander(x, N - 1)
    Vi = np.linalg.inv(V)
    d1 = np.stack([npc.chebval(x, npc.chebder(np.eye(N)[k])) for k in range(N)], 1) @ Vi
    d2 = np.stack([npc.chebval(x, npc.chebder(np.eye(N)[k], 2)) for k in range(N)], 1) @ Vi
    r = ETA / (1 - ETA) + 0.5 * (1 + x)
    L = 4 * d2 + np.diag(1 / r) @ (2 * d1) - m ** 2 * np.diag(1 / r ** 2)
    return L[1:-1, 1:-1]


def eigen_start(N, m, Nm, orc_mod, seed=0):
    """Initial state: conductive T_0, and on mode m the slowest discrete
    eigenvector (zero at the walls) times a complex constant."""
    L = collocation_L(N, m)
    lam, V = np.linalg.eig(L)
    k = np.argmax(lam.real)
    v = np.zeros(N, complex)
    v[1:-1] = V[:, k] / np.abs(V[:, k]).max()
    t, o, u, w = orc_mod.seed_state(N, Nm, ETA, seed, amp=0.0)
    t = t.copy()
    t[m, :] = (0.3 + 0.2j) * v
    return (t, o, u, w), v * (0.3 + 0.2j), lam[k].real


def advance(orc, N, Nm, dt, scheme, state, steps, restart=False):
    run = orc.Run(N, Nm, 0.0, 1.0, dt, scheme, restart=restart)
    run.set_state(*state)
    out = []
    for _ in range(steps):
        run.step(1)
        out.append(run.get_state()["tFR"])
    return out


N_LIN, NM_LIN, M_LIN = 17, 4, 2


@pytest.mark.parametrize("dt", [1e-3, 1e-2])
def test_cn_amplification(orc, dt):
    state, v, lam = eigen_start(N_LIN, M_LIN, NM_LIN, orc)
    z = dt * lam
    g = (1 + z / 2) / (1 - z / 2)
    T = advance(orc, N_LIN, NM_LIN, dt, "CNAB2", state, 4)
    for n, Tn in enumerate(T, 1):
        assert rel(Tn[M_LIN], g ** n * v) < 1e-10


BDF = {  # textbook BDFk: sum_i alpha_i y_{n+1-i} ... as y_{n+1} = (sum a_i y_{n+1-i} + b h f_{n+1})
    "SBDF2": ([4 / 3, -1 / 3], 2 / 3),
    "SBDF3": ([18 / 11, -9 / 11, 2 / 11], 6 / 11),
    "SBDF4": ([48 / 25, -36 / 25, 16 / 25, -3 / 25], 12 / 25),
}


@pytest.mark.parametrize("restart", [True, False])
@pytest.mark.parametrize("scheme", ["SBDF2", "SBDF3", "SBDF4"])
def test_bdf_recurrence(orc, scheme, restart):
    dt = 2e-3
    state, v, lam = eigen_start(N_LIN, M_LIN, NM_LIN, orc)
    a, b = BDF[scheme]
    z = dt * lam
    # restart: constant history; fresh start: zero history
    hist = [v] * len(a) if restart else [v] + [0 * v] * (len(a) - 1)
    T = advance(orc, N_LIN, NM_LIN, dt, scheme, state, 5, restart=restart)
    for n in range(5):
        new = sum(ai * hi for ai, hi in zip(a, hist)) / (1 - b * z)
        hist = [new] + hist[:-1]
        assert rel(T[n][M_LIN], new) < 1e-10, (scheme, restart, n)
    if scheme == "SBDF3" and not restart:
        # the zero-history quirk: first step = (18/11) v / (1 - 6/11 z)
        assert rel(T[0][M_LIN], (18 / 11) * v / (1 - 6 / 11 * z)) < 1e-10


@pytest.mark.parametrize("name", [r[0] for r in RK_SA])
def test_imexrk_amplification(orc, name):
    """Y = 1 + z A' Y with A' the implicit tableau whose stages 2..S are all
    solved with the aD(2,2) operator (diag_diff = .FALSE., diag_index = 2,
    fourierloop_imexrk.F90:108-127); stiffly accurate: y_{n+1} = Y_S.  For a
    uniform diagonal A' = aD and Y_S = 1 + z b^T (I - z aD)^{-1} 1; FW53's
    diagonal is not uniform, and the reference runs it with this A'."""
    dt = 2e-3
    info = orc.scheme_info("IMEXRK", name, "exp" + name)
    A = info["aD"].copy()
    S = len(info["bD"])
    for k in range(1, S):
        A[k, k] = info["aD"][1, 1]
    state, v, lam = eigen_start(N_LIN, M_LIN, NM_LIN, orc)
    z = dt * lam
    R = np.linalg.solve(np.eye(S) - z * A, np.ones(S))[-1]
    if np.allclose(np.diag(info["aD"])[1:], info["aD"][1, 1]):
        b = info["bD"]
        assert abs(R - (1 + z * b @ np.linalg.solve(np.eye(S) - z * info["aD"], np.ones(S)))) < 1e-14
    T = advance(orc, N_LIN, NM_LIN, dt, name, state, 3)
    for n, Tn in enumerate(T, 1):
        assert rel(Tn[M_LIN], R ** n * v) < 1e-10


# ------------------------------------------------------------- conduction
@pytest.mark.parametrize("scheme", ["CNAB2", "SBDF3", "ARS443", "BHR553"])
def test_conduction_is_steady_with_unit_nusselt(orc, scheme):
    N, Nm = 33, 8
    st = orc.seed_state(N, Nm, ETA, 1, amp=0.0)
    run = orc.Run(N, Nm, 1e6, 1.0, 1e-5, scheme, restart=(scheme == "SBDF3"))
    run.set_state(*st)
    run.step(20)
    s = run.get_state()
    d = run.diag()
    assert d["KE"] == 0.0
    assert np.abs(s["urFR"]).max() == 0.0 and np.abs(s["upFR"]).max() == 0.0
    assert rel(s["tFR"], st[0]) < 1e-13
    assert abs(d["Nu_t"] - 1) < 1e-10 and abs(d["Nu_b"] - 1) < 1e-10


# -------------------------------------------------------- temporal order
def final_T(orc, scheme, dt, steps, buo="IMP"):
    N, Nm = 17, 8
    run = orc.Run(N, Nm, 2e4, 1.0, dt, scheme, buo=buo)
    run.set_state(*orc.seed_state(N, Nm, ETA, 5, amp=0.5))
    run.step(steps)
    return run.get_state()["tFR"]


@pytest.mark.parametrize("scheme,order,buo", [("CNAB2", 2, "IMP"), ("ARS443", 3, "IMP"),
                                              ("ARS443", 3, "EXP"), ("BPR353", 3, "IMP"),
                                              ("ARS222", 2, "IMP"), ("CK222", 2, "IMP"), ("PC2", 2, "IMP"),
                                              ("SSP332", 2, "IMP"), ("BPR442", 2, "IMP")])
def test_temporal_order(orc, scheme, order, buo):
    T_end, dt0 = 2e-3, 2e-4
    ref = final_T(orc, scheme, dt0 / 32, 32 * round(T_end / dt0), buo)
    errs = []
    for k in (1, 2, 4):
        errs.append(rel(final_T(orc, scheme, dt0 / k, k * round(T_end / dt0), buo), ref))
    p1 = np.log2(errs[0] / errs[1])
    p2 = np.log2(errs[1] / errs[2])
    assert errs[2] < errs[1] < errs[0]
    assert abs(p2 - order) < 0.35, (errs, p1, p2)


def final_state(orc, scheme, dt, steps, mBC):
    N, Nm = 17, 8
    run = orc.Run(N, Nm, 2e4, 1.0, dt, scheme, mBC=mBC)
    run.set_state(*orc.seed_state(N, Nm, ETA, 5, amp=0.5))
    run.step(steps)
    g = run.get_state()
    return g["tFR"], g["omgFR"], g["upFR"]


@pytest.mark.parametrize("scheme,order,mBC", [("BHR553", 3, "NS"), ("ARS232", 2, "NS"), ("ARS343", 3, "SF")])
def test_assembly_convergence(orc, scheme, order, mBC):
    """Assembly_stage with psi from LAPpsi and Johnston's wall vorticity:
    temperature converges at the scheme's order; vorticity and velocity
    converge too (the wall extrapolation lowers their order)."""
    T_end, dt0 = 2e-3, 2e-4
    ref = final_state(orc, scheme, dt0 / 32, 32 * round(T_end / dt0), mBC)
    errs = np.array([[rel(f, r) for f, r in zip(final_state(orc, scheme, dt0 / k, k * round(T_end / dt0), mBC), ref)]
                     for k in (1, 2, 4)])
    p2 = np.log2(errs[1] / errs[2])
    assert abs(p2[0] - order) < 0.35, (errs, p2)
    assert np.all(errs[2] < errs[1]) and np.all(errs[1] < errs[0])
    assert np.all(p2 > 0.8), p2


def test_buoyancy_flag_only_reaches_imexrk(orc):
    """buo_tscheme is read by the IMEXRK stage only (fourierloop_imexrk.F90:296-301)."""
    assert np.array_equal(final_T(orc, "CNAB2", 2e-4, 5, "IMP"), final_T(orc, "CNAB2", 2e-4, 5, "EXP"))
    assert not np.array_equal(final_T(orc, "ARS443", 2e-4, 5, "IMP"), final_T(orc, "ARS443", 2e-4, 5, "EXP"))
