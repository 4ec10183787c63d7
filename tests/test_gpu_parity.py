"""GPU parity of the time-stepping core against the oracle and the golden
fixtures, through the C-ABI (libfdns.so via its ctypes mirror).

Tolerance: relative Frobenius error <= 1e-10 on the spectral temperature and
streamfunction (and omega, u_r, u_phi where the oracle is run live), and
max relative error <= 1e-10 on the KE and Nusselt series.  The schemes'
timeschemes.F90 names, the IMEX zero-history and restart start-ups, both
velocity boundary conditions, and every radial and azimuthal size tier of
the kernels are covered.
"""
from __future__ import annotations

import ast
import os

import numpy as np
import pytest

from conftest import GOLDEN, rel

pytestmark = pytest.mark.gpu

TOL = 1e-10
ETA = 0.35
FIELDS = ("tFR", "omgFR", "urFR", "upFR", "psi")


def fixture(name):
    z = np.load(os.path.join(GOLDEN, name + ".npz"))
    return ast.literal_eval(str(z["params"])), {k: z[k] for k in ("tFR", "psi", "series")}


def gpu_run(fd, Nr, Nm, Ra, Pr, dt, scheme, restart, mBC, buo, steps, every, seed, solver="auto"):
    g = fd.Run(Nr, Nm, Ra, Pr, dt, scheme, mBC=mBC, buo=buo, restart=restart, n_KE=every, solver=solver)
    g.seed(seed)
    g.step(steps)
    out = g.get_state()
    out["series"] = g.series()
    out["diag"] = g.diag()
    g.close()
    return out


def series_err(a, b):
    a, b = np.asarray(a), np.asarray(b)
    assert a.shape == b.shape, (a.shape, b.shape)
    err = 0.0
    for c in (1, 2, 3):          # KE, Nu_t, Nu_b; t is checked exactly below
        err = max(err, float(np.max(np.abs(a[:, c] - b[:, c]) / np.maximum(np.abs(b[:, c]), 1e-300))))
    assert np.allclose(a[:, 0], b[:, 0], rtol=1e-14, atol=0)
    return err


SMALL = ["s_cnab2", "s_cnab3", "s_sbdf2", "s_sbdf3", "s_sbdf3_fresh", "s_sbdf4", "s_ars443", "s_ars443_exp",
         "s_bpr353", "s_cnab2_sf", "s_ars443_sf"]


@pytest.mark.parametrize("solver", ["inverse", "lu"])
@pytest.mark.parametrize("name", SMALL)
def test_scheme_fixture(fd, name, solver):
    case, want = fixture(name)
    got = gpu_run(fd, *case, solver=solver)
    print(name, solver, {k: f"{rel(got[k], want[k]):.2e}" for k in ("tFR", "psi")}, flush=True)
    assert rel(got["tFR"], want["tFR"]) <= TOL
    assert rel(got["psi"], want["psi"]) <= TOL
    assert series_err(got["series"], want["series"]) <= TOL


FAMILIES = ["s_ars222", "s_bhr553", "s_kc343_sf", "s_smr333", "s_cb3e", "s_rk4"]


@pytest.mark.parametrize("name", FAMILIES)
def test_family_fixture(fd, name):
    """The other IMEXRK families (stiffly accurate, Assembly_stage, own
    diagonals incl. CB3e's zero-sum column) and RK4 against their fixtures,
    the solver each picks (LU for the own-diagonal schemes)."""
    case, want = fixture(name)
    got = gpu_run(fd, *case)
    print(name, {k: f"{rel(got[k], want[k]):.2e}" for k in ("tFR", "psi")}, flush=True)
    assert rel(got["tFR"], want["tFR"]) <= TOL
    assert rel(got["psi"], want["psi"]) <= TOL
    assert series_err(got["series"], want["series"]) <= TOL


@pytest.mark.parametrize("solver", ["inverse", "lu"])
def test_c0_full_trajectory(fd, solver):
    """The reference's CPU case: CNAB2, Nr=65, Nm=128, Ra=1e6, dt=1e-6,
    1000 steps, (t, KE, Nu_t, Nu_b) every 10 steps."""
    case, want = fixture("c0_cnab2_nr65_nm128")
    got = gpu_run(fd, *case, solver=solver)
    eT, eP, eS = rel(got["tFR"], want["tFR"]), rel(got["psi"], want["psi"]), series_err(got["series"], want["series"])
    print(f"c0 {solver}: T {eT:.2e} psi {eP:.2e} series {eS:.2e}", flush=True)
    assert eT <= TOL and eP <= TOL and eS <= TOL


def live(fd, orc, Nr, Nm, Ra, Pr, dt, scheme, steps, seed, restart=False, mBC="NS", buo="IMP", threads=0,
         solver="auto"):
    g = fd.Run(Nr, Nm, Ra, Pr, dt, scheme, mBC=mBC, buo=buo, restart=restart, solver=solver)
    g.seed(seed)
    g.step(steps)
    a = g.get_state()
    da = g.diag()
    g.close()
    o = orc.Run(Nr, Nm, Ra, Pr, dt, scheme, mBC=mBC, buo=buo, restart=restart, nthreads=threads)
    o.set_state(*orc.seed_state(Nr, Nm, ETA, seed))
    o.step(steps)
    b = o.get_state()
    db = o.diag()
    errs = {k: rel(a[k], b[k]) for k in FIELDS}
    print(f"{scheme} Nr={Nr} Nm={Nm} steps={steps} solver={solver}:", {k: f"{v:.2e}" for k, v in errs.items()},
          flush=True)
    return errs, da, db


# every radial tier (M = 2Nr-2 up to 128, 256, 512, 1024) and azimuthal
# tier (Np = 3Nm up to 768, 2304, 4608) of the kernels
TIERS = [
    (17, 8, 1e4, 1.0, 1e-4, "CNAB2", 10, 1),
    (65, 64, 1e6, 1.0, 1e-6, "SBDF3", 10, 2),
    (97, 32, 1e6, 0.7, 1e-6, "ARS443", 4, 3),
    (129, 32, 1e6, 1.0, 1e-6, "SBDF2", 10, 4),
    (257, 16, 1e7, 1.0, 1e-7, "ARS443", 3, 5),
    (257, 16, 1e7, 7.0, 1e-7, "CNAB2", 5, 6),
    (513, 8, 1e9, 1.0, 2e-8, "BPR353", 2, 7),
    (33, 512, 1e6, 1.0, 1e-6, "CNAB2", 5, 8),
    (33, 768, 1e6, 1.0, 1e-6, "BPR353", 2, 9),
    (17, 1536, 1e6, 1.0, 1e-6, "ARS443", 2, 10),
]


@pytest.mark.parametrize("case", TIERS, ids=[f"{c[5]}-{c[0]}x{c[1]}" for c in TIERS])
def test_size_tiers_vs_oracle(fd, orc, case):
    Nr, Nm, Ra, Pr, dt, scheme, steps, seed = case
    errs, da, db = live(fd, orc, Nr, Nm, Ra, Pr, dt, scheme, steps, seed, restart=(scheme.startswith("SBDF")))
    assert max(errs.values()) <= TOL
    for k in ("KE", "Nu_t", "Nu_b"):
        assert abs(da[k] - db[k]) <= TOL * max(abs(db[k]), 1e-300)


@pytest.mark.parametrize("case", [(65, 64, 1e6, 1.0, 1e-6, "SBDF3", 10, 2), (97, 32, 1e6, 0.7, 1e-6, "ARS443", 4, 3),
                                  (129, 32, 1e6, 1.0, 1e-6, "CNAB2", 10, 4)], ids=["65", "97", "129"])
def test_lu_mode_vs_oracle(fd, orc, case):
    """The dgetrs-order solves at the sizes where auto picks the inverses."""
    Nr, Nm, Ra, Pr, dt, scheme, steps, seed = case
    errs, da, db = live(fd, orc, Nr, Nm, Ra, Pr, dt, scheme, steps, seed, restart=(scheme.startswith("SBDF")),
                        solver="lu")
    assert max(errs.values()) <= TOL
    for k in ("KE", "Nu_t", "Nu_b"):
        assert abs(da[k] - db[k]) <= TOL * max(abs(db[k]), 1e-300)


@pytest.mark.parametrize("scheme", ["ARS222", "PC2", "SSP332", "SSP543", "FW53", "CK222", "BPR442"])
def test_imexrk_tableaux_vs_oracle(fd, orc, scheme):
    """Every stiffly accurate IMEXRK tableau of the registry, both solve modes."""
    for solver in ("inverse", "lu"):
        errs, da, db = live(fd, orc, 33, 16, 1e6, 1.0, 1e-5, scheme, 4, 21, solver=solver)
        assert max(errs.values()) <= TOL, (scheme, solver)


@pytest.mark.parametrize("scheme", ["EUL111", "ARS232", "ARS233", "ARS343", "MARS343", "IJ3", "PR433", "DBM453",
                                    "KC343", "KC564", "KC774", "KC885", "CFN343", "BHR553", "CFN564"])
def test_assembly_schemes_vs_oracle(fd, orc, scheme):
    """IMEXRK schemes whose steps end with Assembly_stage (LAPpsi, Johnston's
    wall vorticity), both solve modes; no slip and stress free."""
    for solver, mBC in (("inverse", "NS"), ("lu", "NS"), ("inverse", "SF")):
        errs, da, db = live(fd, orc, 33, 16, 1e6, 1.0, 1e-5, scheme, 4, 22, solver=solver, mBC=mBC)
        assert max(errs.values()) <= TOL, (scheme, solver, mBC)
        for k in ("KE", "Nu_t", "Nu_b"):
            assert abs(da[k] - db[k]) <= TOL * max(abs(db[k]), 1e-300), (scheme, k)


@pytest.mark.parametrize("scheme", ["SMR333", "CB3e", "CB3f", "CB4", "LZ3L1", "LZ4A1"])
def test_own_diagonal_schemes_vs_oracle(fd, orc, scheme):
    """IMEXRK schemes that solve each stage with its own diagonal's operator
    (LU solves); no slip and stress free.  CB3e's third explicit column sums
    to zero (aA(4,3) + bA(3) = 0) while its entries do not: the nonlinear pass
    of that stage is skipped and its explicit terms are built from the
    products of the pass before (steptime_imexrk.F90:153-166)."""
    for mBC in ("NS", "SF"):
        errs, da, db = live(fd, orc, 33, 16, 1e6, 1.0, 1e-5, scheme, 4, 23, mBC=mBC)
        assert max(errs.values()) <= TOL, (scheme, mBC)
    with pytest.raises(fd.FdnsError):
        fd.Run(33, 16, 1e6, 1.0, 1e-5, scheme, solver="inverse")


def test_zero_diagonal_scheme_refused(fd):
    """LZ3A4a's stages 2 and 4 have aD(k,k) = 0, for which the
    vorticity-streamfunction operator is singular: refused at creation."""
    with pytest.raises(fd.FdnsError, match="singular"):
        fd.Run(33, 16, 1e6, 1.0, 1e-5, ("IMEXRK", "LZ3A4a", "expLZ3A4a"))


def test_c1_full_size(fd, orc):
    """ARS443, Nr=257, Nm=768, Ra=1e8, dt=1e-7: two steps at the full size."""
    errs, da, db = live(fd, orc, 257, 768, 1e8, 1.0, 1e-7, "ARS443", 2, 1000, threads=min(16, os.cpu_count() or 1))
    assert max(errs.values()) <= TOL


def test_c1_full_size_assembly_scheme(fd, orc):
    """BHR553 (Assembly_stage: LAPpsi psi solve, Johnston walls), Nr=257,
    Nm=768, Ra=1e8, dt=1e-7: one step at the full size, LU solves."""
    errs, da, db = live(fd, orc, 257, 768, 1e8, 1.0, 1e-7, "BHR553", 1, 1001, threads=min(16, os.cpu_count() or 1))
    assert max(errs.values()) <= TOL


@pytest.mark.parametrize("k", [0, 1, 2])
def test_c2_members(fd, orc, k):
    """Three members of the 64-run ensemble (one per scheme), 10 steps."""
    import sys
    from conftest import REPO
    sys.path.insert(0, REPO)
    from bench import ensemble_members
    m = ensemble_members()[k]
    errs, _, _ = live(fd, orc, m["Nr"], m["Nm"], m["Ra"], m["Pr"], m["dt"], m["scheme"], 10, m["seed"],
                      restart=m["restart"], threads=min(16, os.cpu_count() or 1))
    assert max(errs.values()) <= TOL


def test_c4_full_size_one_step(fd, orc):
    """BPR353 at Nr=513, Nm=1536, Ra=1e9, dt=2e-8 (the largest tiles and
    transforms, LU solves), one step with the nonlinear products of a seeded
    perturbation, against the oracle."""
    errs, da, db = live(fd, orc, 513, 1536, 1e9, 1.0, 2e-8, "BPR353", 1, 4001, threads=min(16, os.cpu_count() or 1))
    assert max(errs.values()) <= TOL


def test_c4_full_size_properties(fd, orc):
    """BPR353 at Nr=513, Nm=1536 (the largest tiles): (1) pure conduction
    stays put with u = 0 and Nu = 1; (2) with Ra = 0 every mode evolves alone,
    so modes 0..8 of the full-size run equal an oracle run with Nm = 8 (the
    seeded perturbation of mode m does not depend on Nm)."""
    Nr, Nm, dt = 513, 1536, 2e-8
    g = fd.Run(Nr, Nm, 1e9, 1.0, dt, "BPR353")
    g.seed(4000, amp=0.0)
    s0 = g.get_state()
    g.step(2)
    s = g.get_state()
    d = g.diag()
    g.close()
    assert not np.any(s["urFR"]) and not np.any(s["upFR"])
    assert rel(s["tFR"], s0["tFR"]) < 1e-11
    assert abs(d["Nu_t"] - 1) < TOL and abs(d["Nu_b"] - 1) < TOL and d["KE"] == 0.0

    g = fd.Run(Nr, Nm, 0.0, 1.0, dt, "BPR353")
    g.seed(4000)
    g.step(2)
    T = g.get_state()["tFR"]
    g.close()
    assert np.all(np.isfinite(T))
    o = orc.Run(Nr, 8, 0.0, 1.0, dt, "BPR353")
    o.set_state(*orc.seed_state(Nr, 8, ETA, 4000))
    o.step(2)
    Tref = o.get_state()["tFR"]
    e = rel(T[:9], Tref)
    print(f"c4 modes 0..8 vs oracle: {e:.2e}", flush=True)
    assert e <= TOL
