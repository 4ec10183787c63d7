"""writeKE_spectral's power budget, mean-flow energy and mean-flow balance,
and calculate_spectra, on the GPU against the oracle, from a developed flow
(mean flow of the order of the other modes) and every scheme family; in the
graph-replayed steps of n_KE = 3, in variable-dt runs, and the refusals.
"""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

Nr, Nm = 33, 16
STATE = ("tFR", "omgFR", "urFR", "upFR")
TOL = 1e-10


@pytest.fixture(scope="module")
def developed(orc):
    A = orc.Run(Nr, Nm, 1e6, 1.0, 1e-5, "CNAB2")
    A.set_state(*orc.seed_state(Nr, Nm, 0.35, 3, amp=1e-1))
    A.step(200)
    g = A.get_state()
    return tuple(g[f] for f in STATE)


def _compare(g, o, tag):
    bg, ug = g.budget()
    bo, uo = o.budget()
    e = {k: abs(bg[k] - bo[k]) / abs(bo[k]) for k in ("KE_mean", "vis", "buo")}
    e["t"] = abs(bg["t"] - bo["t"])
    for c, name in enumerate(("du/dt", "advection", "viscous")):
        e[name] = np.abs(ug[:, c] - uo[:, c]).max() / np.abs(uo[:, c]).max()
    kg, tg = g.spectra()
    ko, to = o.spectra()
    e["KEspec"] = np.abs(kg - ko).max() / np.abs(ko).max()
    e["Tspec"] = np.abs(tg - to).max() / np.abs(to).max()
    print(tag, {k: f"{v:.1e}" for k, v in e.items()}, flush=True)
    assert e.pop("t") < 1e-15
    assert max(e.values()) < TOL, tag


@pytest.mark.parametrize("scheme,mBC,dt", [("CNAB2", "NS", 1e-6), ("SBDF3", "SF", 1e-6), ("ARS443", "NS", 1e-6),
                                           ("BHR553", "SF", 1e-6), ("CB3f", "NS", 1e-6), ("RK4", "NS", 1e-7)])
def test_budget_vs_oracle(fd, orc, developed, scheme, mBC, dt):
    g = fd.Run(Nr, Nm, 1e6, 1.0, dt, scheme, mBC=mBC, n_KE=3)
    g.set_state(*developed)
    o = orc.Run(Nr, Nm, 1e6, 1.0, dt, scheme, mBC=mBC)
    o.set_state(*developed)
    for steps in (1, 4):          # a recorded first step, then graph-replayed steps
        g.step(steps)
        o.step(steps)
        _compare(g, o, f"{scheme} {mBC} after {steps}")


def test_budget_variable_dt(fd, orc, developed):
    kw = dict(vardt=True, CFL=0.5, dt_coef=1.0, dt_max=1e-2)
    g = fd.Run(Nr, Nm, 1e6, 1.0, 1e-5, "SBDF3", restart=True, n_KE=1, **kw)
    g.set_state(*developed)
    o = orc.Run(Nr, Nm, 1e6, 1.0, 1e-5, "SBDF3", restart=True, **kw)
    o.set_state(*developed)
    g.step(3)
    o.step(3)
    _compare(g, o, "SBDF3 variable dt")


def test_budget_needs_recording_run(fd, developed):
    g = fd.Run(Nr, Nm, 1e6, 1.0, 1e-6, "CNAB2")
    g.set_state(*developed)
    g.step(1)
    with pytest.raises(fd.FdnsError):
        g.budget()
    h = fd.Run(Nr, Nm, 1e6, 1.0, 1e-6, "CNAB2", n_KE=1)
    h.set_state(*developed)
    with pytest.raises(fd.FdnsError):
        h.budget()                 # before the first step
    h.step(1)
    h.budget()


def test_driver_budget_files(fd, tmp_path):
    """The driver's power_, upbalance_, KEspec_ and Temp_spec_ files and the
    KE file's sixth column against the library run of the same input."""
    import subprocess
    from test_driver import driver, write_input
    inp = write_input(tmp_path / "input.txt", Nr=Nr, Nm=Nm, Ra=1e6, steps=6, dt=1e-5, every=3, seed=9, tag="b")
    inp.write_text(inp.read_text().replace("n_KEspec=100000", "n_KEspec=2"))
    p = subprocess.run([driver(), str(inp)], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stderr
    g = fd.Run(Nr, Nm, 1e6, 1.0, 1e-5, "CNAB2", n_KE=3)
    g.seed(9, amp=0.001)
    rows_b, rows_u, rows_k, rows_t = [], [], [], []
    for k in range(1, 7):
        g.step(1)
        if k % 3 == 0:
            b, ub = g.budget()
            rows_b.append((b["t"], b["vis"], b["buo"], b["KE_mean"]))
            rows_u.append(ub)
        if k % 2 == 0:
            ks, ts = g.spectra()
            rows_k.append(ks)
            rows_t.append(ts)
    pw = np.loadtxt(tmp_path / "power_b", ndmin=2)
    ke = np.loadtxt(tmp_path / "KE_b", ndmin=2)
    up = np.loadtxt(tmp_path / "upbalance_b", ndmin=2)
    ksf = np.loadtxt(tmp_path / "KEspec_b", ndmin=2)
    tsf = np.loadtxt(tmp_path / "Temp_spec_b", ndmin=2)
    rb = np.array(rows_b)
    # the files carry 13 significant digits (es20.12)
    assert pw.shape == (2, 3) and np.allclose(pw, rb[:, :3], rtol=1e-11, atol=0)
    assert ke.shape == (2, 6) and np.allclose(ke[:, 5], rb[:, 3], rtol=1e-11, atol=1e-300)
    assert up.shape == (2 * Nr, 4)
    assert np.allclose(up[:, 1:], np.concatenate(rows_u), rtol=1e-11, atol=1e-12 * np.abs(up[:, 1:]).max())
    assert ksf.shape == (3 * (Nm + 1), 3) and tsf.shape == (3 * (Nm + 1), 3)
    assert np.array_equal(ksf[:, 0], np.repeat([2, 4, 6], Nm + 1)) and np.array_equal(ksf[:, 1], np.tile(np.arange(Nm + 1), 3))
    assert np.allclose(ksf[:, 2], np.concatenate(rows_k), rtol=1e-11, atol=0)
    assert np.allclose(tsf[:, 2], np.concatenate(rows_t), rtol=1e-11, atol=0)


def test_budget_after_ensemble_steps(fd, developed):
    """Runs that record (n_KE > 0) keep the budget inputs inside batched
    steps too: after the ensemble closes, each member's budget equals the
    one of the same run stepped alone."""
    def make(scheme):
        r = fd.Run(Nr, Nm, 1e6, 1.0, 1e-6, scheme, n_KE=3)
        r.set_state(*developed)
        return r
    schemes = ("CNAB2", "SBDF3", "ARS443", "BHR553")
    runs = [make(s) for s in schemes]
    ens = fd.Ensemble(runs)
    ens.step(5)
    ens.close()
    for s, r in zip(schemes, runs):
        solo = make(s)
        solo.step(5)
        (bg, ug), (bs, us) = r.budget(), solo.budget()
        assert all(bg[k] == bs[k] for k in bg), s
        assert np.array_equal(ug, us), s
