"""The explicit RK path (time_scheme_type = 'RK', RK2 / RK4) on the GPU
against the oracle: fresh starts and restarts, no slip and stress free, the
KE / Nusselt diagnostics, the checkpoint and snapshot records of
store_checkpoint / store_snapshot, and the limits (no variable dt, no
ensembles or pools).
"""
from __future__ import annotations

import numpy as np
import pytest

from conftest import rel
from test_gpu_parity import FIELDS, TOL, live

pytestmark = pytest.mark.gpu

STATE = ("tFR", "omgFR", "urFR", "upFR")


@pytest.mark.parametrize("scheme", ["RK2", "RK4"])
@pytest.mark.parametrize("mBC", ["NS", "SF"])
@pytest.mark.parametrize("restart", [False, True])
def test_rk_vs_oracle(fd, orc, scheme, mBC, restart):
    errs, da, db = live(fd, orc, 33, 16, 1e5, 1.0, 2e-6, scheme, 6, 31, restart=restart, mBC=mBC)
    assert max(errs.values()) <= TOL, (scheme, mBC, restart)
    for k in ("KE", "Nu_t", "Nu_b"):
        assert abs(da[k] - db[k]) <= TOL * max(abs(db[k]), 1e-300), (scheme, k)


def test_rk_larger_shape_vs_oracle(fd, orc):
    """Radial tier M = 256 and Np = 3 Nm = 192."""
    errs, _, _ = live(fd, orc, 129, 64, 1e6, 1.0, 2e-8, "RK4", 3, 32, threads=8)
    assert max(errs.values()) <= TOL


def test_rk_step_by_step_equals_batched_steps(fd):
    """One step per call (no graph) against four steps in one call."""
    a = fd.Run(33, 16, 1e5, 1.0, 2e-6, "RK4")
    b = fd.Run(33, 16, 1e5, 1.0, 2e-6, "RK4")
    a.seed(4)
    b.seed(4)
    for _ in range(4):
        a.step(1)
    b.step(4)
    sa, sb = a.get_state(), b.get_state()
    for f in FIELDS:
        assert np.array_equal(sa[f], sb[f]), f


def test_rk_files(fd, tmp_path):
    from fdns import files
    g = fd.Run(33, 16, 1e5, 1.0, 2e-6, "RK4")
    g.seed(5)
    g.step(3)
    g.write_checkpoint(tmp_path / "checkpoint_00001")
    c = files.read_checkpoint(tmp_path / "checkpoint_00001")
    assert (c["n_imp"], c["n_exp"], c["n_max"]) == (4, 4, 4)
    for k in ("imp_temp", "exp_temp", "imp_vort", "exp_vort"):
        assert c[k].shape == (4, 17, 33) and not c[k].any()
    st = g.get_state()
    for f in STATE:
        assert np.array_equal(c[f], st[f])
    g.write_snapshot(tmp_path / "snapshot_plot_00001")
    s = files.read_snapshot(tmp_path / "snapshot_plot_00001")
    assert "omgFR_check" not in s           # store_snapshot, not the IMEXRK variant


def test_rk_variable_dt_vs_oracle(fd, orc):
    """From a moving fluid, RK4 with l_vartimestep: each step's dt (the
    CFL and diffusive limits of the previous step's last pass) and the state
    after 6 steps against the oracle."""
    A = orc.Run(33, 16, 1e6, 1.0, 1e-5, "CNAB2")
    A.set_state(*orc.seed_state(33, 16, 0.35, 3, amp=1e-1))
    A.step(100)
    g = A.get_state()
    st = tuple(g[f] for f in STATE)
    kw = dict(vardt=True, CFL=0.5, dt_coef=1.0, dt_max=1e-2)
    r = fd.Run(33, 16, 1e6, 1.0, 1e-6, "RK4", **kw)
    r.set_state(*st)
    o = orc.Run(33, 16, 1e6, 1.0, 1e-6, "RK4", **kw)
    o.set_state(*st)
    dg, do = [], []
    for _ in range(6):
        r.step(1)
        o.step(1)
        dg.append(r.dt())
        do.append(o.dt())
    a, b = r.get_state(), o.get_state()
    errs = {f: rel(a[f], b[f]) for f in STATE}
    e_dt = max(abs(x - y) / y for x, y in zip(dg, do))
    print("RK4 variable dt:", [f"{d:.3e}" for d in dg], f"dt err {e_dt:.1e}", {k: f"{v:.1e}" for k, v in errs.items()},
          flush=True)
    assert dg[0] == 1e-6 and dg[1] != 1e-6
    assert e_dt < 1e-12
    assert max(errs.values()) < TOL


def test_rk_limits(fd):
    with pytest.raises(fd.FdnsError):
        fd.Run(33, 16, 1e5, 1.0, 2e-6, ("RK", "CN", "RK4"))


def test_rk_restart_continues_run(fd, orc):
    """A restart from one state continues the run (the RK path keeps no
    history), and matches the oracle restarted from the same state."""
    A = fd.Run(33, 16, 1e5, 1.0, 2e-6, "RK4")
    A.seed(6)
    A.step(3)
    g = A.get_state()
    st = tuple(g[f] for f in STATE)
    t = A.diag()["t"]
    A.step(3)
    ref = A.get_state()
    B = fd.Run(33, 16, 1e5, 1.0, 2e-6, "RK4", restart=True)
    B.restart([st], t)
    B.step(3)
    got = B.get_state()
    o = orc.Run(33, 16, 1e5, 1.0, 2e-6, "RK4", restart=True)
    o.restart([st], t)
    o.step(3)
    want = o.get_state()
    assert max(rel(got[f], want[f]) for f in STATE) < TOL
    # the restart's first stage takes upFC = u_phi's values (init.F90:395),
    # not the coefficients the uninterrupted run holds; that changes only the
    # first stage's mean-flow forcing, far below the fields here
    assert max(rel(got[f], ref[f]) for f in STATE) < 1e-9


def test_rk_developed_flow_vs_oracle(fd, orc):
    """From a developed flow (200 CNAB2 steps of the oracle at Ra = 1e6), whose
    mean flow is of the order of the other modes: the mean-flow right-hand
    side of the RK path (Fourier-space sums over the modes) carries weight."""
    A = orc.Run(33, 16, 1e6, 1.0, 1e-5, "CNAB2")
    A.set_state(*orc.seed_state(33, 16, 0.35, 3, amp=1e-1))
    A.step(200)
    g = A.get_state()
    st = tuple(g[f] for f in STATE)
    assert np.abs(g["upFR"][0]).max() > 0.1 * np.abs(g["upFR"]).max()
    for scheme in ("RK2", "RK4"):
        r = fd.Run(33, 16, 1e6, 1.0, 1e-6, scheme)
        r.set_state(*st)
        r.step(5)
        o = orc.Run(33, 16, 1e6, 1.0, 1e-6, scheme)
        o.set_state(*st)
        o.step(5)
        a, b = r.get_state(), o.get_state()
        errs = {f: rel(a[f], b[f]) for f in STATE}
        errs["mean flow"] = rel(a["upFR"][0], b["upFR"][0])
        print(scheme, "developed flow:", {k: f"{v:.2e}" for k, v in errs.items()}, flush=True)
        assert max(errs.values()) < TOL, scheme


def test_rk_in_ensemble_and_pool_equal_solo(fd):
    """RK2 and RK4 runs batched with an ARS443 run (ensemble, graph replays)
    and an RK4 job in a pool: bitwise equal to their solo runs."""
    def make(scheme, seed):
        r = fd.Run(33, 16, 1e5, 1.0, 2e-6, scheme)
        r.seed(seed)
        return r
    members = [("RK2", 41), ("RK4", 42), ("RK4", 43), ("ARS443", 44)]
    runs = [make(s, k) for s, k in members]
    ens = fd.Ensemble(runs)
    ens.step(5)
    got = [r.get_state() for r in runs]
    ens.close()
    for (s, k), g in zip(members, got):
        solo = make(s, k)
        solo.step(5)
        want = solo.get_state()
        assert all(np.array_equal(g[f], want[f]) for f in STATE), s
    P = fd.Pool(1 << 30)
    jid = P.submit(33, 16, 1e5, 1.0, 2e-6, "RK4", 4, seed=45)
    while P.poll(jid)[0] != 2:
        P.step()
    res = P.result(jid)
    P.close()
    solo = make("RK4", 45)
    solo.step(4)
    want = solo.get_state()
    assert all(np.array_equal(res[f], want[f]) for f in STATE)
