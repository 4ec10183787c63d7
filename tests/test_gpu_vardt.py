"""Variable dt (l_vartimestep) on the GPU against the oracle.

From a moving fluid the CFL limit decides every step's dt: the dt sequence
and the state after the steps match the oracle's; with dt below the limit the
run is the constant-dt run bit for bit; the Fortran driver writes each row's
dt in the KE file; ensembles and pools refuse variable-dt runs.
"""
from __future__ import annotations

import subprocess

import numpy as np
import pytest

from conftest import rel

pytestmark = pytest.mark.gpu

Nr, Nm, ETA = 33, 16, 0.35
FIELDS = ("tFR", "omgFR", "urFR", "upFR")
TOL = 1e-10


@pytest.fixture(scope="module")
def moving(orc):
    """A state with a developed flow (oracle, constant dt)."""
    A = orc.Run(Nr, Nm, 1e6, 1.0, 1e-5, "CNAB2", eta=ETA)
    A.set_state(*orc.seed_state(Nr, Nm, ETA, 3, amp=1e-1))
    A.step(200)
    g = A.get_state()
    return tuple(g[f] for f in FIELDS)


CASES = [("CNAB2", False, False), ("CNAB2", False, True), ("SBDF3", True, False), ("SBDF4", True, False),
         ("ARS443", False, False), ("BPR353", True, False), ("BHR553", False, False), ("CB3e", False, False)]


@pytest.mark.parametrize("scheme,restart,optimize", CASES,
                         ids=[f"{c[0]}{'-restart' if c[1] else ''}{'-opt' if c[2] else ''}" for c in CASES])
def test_vardt_vs_oracle(fd, orc, moving, scheme, restart, optimize):
    kw = dict(vardt=True, CFL=0.5, dt_coef=1.0, dt_max=1e-2, optimizedt=optimize)
    g = fd.Run(Nr, Nm, 1e6, 1.0, 1e-5, scheme, restart=restart, **kw)
    g.set_state(*moving)
    o = orc.Run(Nr, Nm, 1e6, 1.0, 1e-5, scheme, eta=ETA, restart=restart, **kw)
    o.set_state(*moving)
    dg, do = [], []
    for _ in range(8):
        g.step(1)
        o.step(1)
        dg.append(g.dt())
        do.append(o.dt())
    e_dt = max(abs(a - b) / b for a, b in zip(dg, do))
    sg, so = g.get_state(), o.get_state()
    errs = {f: rel(sg[f], so[f]) for f in FIELDS}
    print(f"{scheme} restart={restart} opt={optimize}: dt {dg[0]:.3e}..{dg[-1]:.3e} dt err {e_dt:.1e}",
          {k: f"{v:.2e}" for k, v in errs.items()}, flush=True)
    assert e_dt < 1e-12
    assert len(set(dg)) > 2                       # dt moved
    assert max(errs.values()) < TOL
    assert abs(g.diag()["t"] - o.diag()["t"]) < 1e-15


def test_vardt_capped_equals_constant(fd, orc):
    """dt = dt_max below the CFL limit: no dt change, so the variable-dt run
    (LU solves) equals the constant-dt LU run bit for bit."""
    st = orc.seed_state(Nr, Nm, ETA, 4, amp=1e-2)
    for scheme in ("CNAB2", "SBDF3", "ARS443"):
        a = fd.Run(Nr, Nm, 1e5, 1.0, 1e-6, scheme, vardt=True, CFL=0.5, dt_coef=1.0, dt_max=1e-6)
        b = fd.Run(Nr, Nm, 1e5, 1.0, 1e-6, scheme, solver="lu")
        a.set_state(*st)
        b.set_state(*st)
        a.step(5)
        b.step(5)
        assert a.dt() == 1e-6
        sa, sb = a.get_state(), b.get_state()
        for f in FIELDS:
            assert np.array_equal(sa[f], sb[f]), (scheme, f)


def test_vardt_runs_step_alone(fd):
    r = fd.Run(Nr, Nm, 1e5, 1.0, 1e-6, "CNAB2", vardt=True, CFL=0.5, dt_coef=1.0, dt_max=1e-4)
    r.seed(1)
    with pytest.raises(fd.FdnsError):
        fd.Ensemble([r])
    P = fd.Pool(1 << 30)
    with pytest.raises(fd.FdnsError):
        P.submit(Nr, Nm, 1e5, 1.0, 1e-6, "CNAB2", 4, seed=1, vardt=True, CFL=0.5, dt_coef=1.0, dt_max=1e-4)
    P.close()


def test_driver_vardt_restart_rows(fd, tmp_path):
    """The driver restarted (one checkpoint) from a moving fluid with
    l_vartimestep: the KE rows carry each step's CFL-limited dt and match the
    library run restarted from the same state."""
    from test_driver import driver, write_input
    a = fd.Run(Nr, Nm, 1e6, 1.0, 1e-5, "CNAB2")
    a.seed(9, amp=0.1)
    a.step(200)
    a.write_checkpoint(tmp_path / "checkpoint_00003")
    g = a.get_state()
    t0 = a.diag()["t"]
    inp = write_input(tmp_path / "input.txt", Nr=Nr, Nm=Nm, Ra=1e6, steps=12, dt=1e-5, vardt="TRUE", every=1)
    text = inp.read_text().replace("CFL=0.1", "CFL=0.5").replace("l_restart=.FALSE.", "l_restart=.TRUE.")
    text = text.replace("n_restart=0", "n_restart=200").replace("n_restart_point=1", "n_restart_point=3")
    inp.write_text(text)
    p = subprocess.run([driver(), str(inp)], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stderr
    ke = np.loadtxt(tmp_path / "KE_t", ndmin=2)
    r = fd.Run(Nr, Nm, 1e6, 1.0, 1e-5, "CNAB2", restart=True, vardt=True, CFL=0.5, dt_coef=1.0, dt_max=1e-2)
    r.restart([tuple(g[f] for f in FIELDS)], t0)
    rows = []
    for _ in range(12):
        r.step(1)
        d = r.diag()
        rows.append((r.dt(), d["t"], d["KE"]))
    rows = np.array(rows)
    assert ke.shape == (12, 6)
    # the files carry 13 significant digits (es20.12)
    assert np.allclose(ke[:, :3], rows, rtol=1e-11, atol=0)
    assert len(set(ke[:, 0])) > 2 and ke[:, 0].max() < 1e-2
