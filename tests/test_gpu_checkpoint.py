"""Checkpoints, snapshots and restarts on the GPU.

* fdns_write_checkpoint / fdns_write_snapshot write the reference's files
  (output.F90:359-516): header and fields of the run, and history arrays equal
  to the oracle's after the same steps (to the parity tolerance).
* Their bytes equal those of the Fortran driver's own unformatted writes of
  the same run (the reference writes these files with Fortran I/O).
* fdns_restart from the states of consecutive checkpoints continues the
  uninterrupted run, and matches the oracle restarted from the same states.
* The Fortran driver restarted from its own checkpoint files continues its
  KE / Nusselt series.
"""
from __future__ import annotations

import os
import subprocess

import numpy as np
import pytest

from conftest import rel

pytestmark = pytest.mark.gpu

Nr, Nm, ETA, SEED = 33, 16, 0.35, 5
FIELDS = ("tFR", "omgFR", "urFR", "upFR")
HIST = ("imp_temp", "exp_temp", "imp_vort", "exp_vort", "imp_ub", "exp_ub")
TOL = 1e-10        # relative, as every GPU-vs-oracle parity check


def _rel(a, b):
    b = np.asarray(b)
    n = np.abs(b).max()
    return np.abs(np.asarray(a) - b).max() / n if n > 0 else np.abs(a).max()


@pytest.mark.parametrize("scheme", ["CNAB2", "CNAB3", "SBDF3", "SBDF4", "ARS443"])
def test_checkpoint_file_vs_oracle(fd, orc, tmp_path, scheme):
    from fdns import files
    S, dt = 7, 1e-5
    g = fd.Run(Nr, Nm, 1e5, 1.0, dt, scheme)
    g.seed(SEED)
    g.step(S)
    path = tmp_path / "checkpoint_00001"
    g.write_checkpoint(path)
    c = files.read_checkpoint(path)
    ni, ne = g.orders()
    assert (c["Nm"], c["Nr"], c["n_imp"], c["n_exp"], c["n_max"]) == (Nm, Nr, ni, ne, max(ni, ne))
    assert c["dt"] == dt and np.array_equal(c["dt_array"], [dt] * max(ni, ne))
    assert c["time"] == g.diag()["t"]
    st = g.get_state()
    for f in FIELDS:
        assert np.array_equal(c[f], st[f])
    o = orc.Run(Nr, Nm, 1e5, 1.0, dt, scheme)
    o.set_state(*orc.seed_state(Nr, Nm, ETA, SEED))
    o.step(S)
    imex = fd.SCHEMES[scheme][0] == "IMEX"
    ho = o.history(ni if imex else 1, ne if imex else 1)
    so = o.get_state()
    # the mean-flow histories are measured on the scale of the field they
    # belong to, as the state parity measures u_phi over all modes: the m = 0
    # mean flow is O(amp^2) below the m >= 1 modes early in a run
    scale = {"imp_ub": np.abs(so["upFR"]).max(), "exp_ub": np.abs(so["urFR"]).max() * np.abs(so["omgFR"]).max()}
    errs = {k: (np.abs(c[k] - ho[k]).max() / scale[k] if k in scale else _rel(c[k], ho[k])) for k in HIST}
    print(scheme, "checkpoint histories vs oracle:", {k: f"{v:.2e}" for k, v in errs.items()}, flush=True)
    for k in HIST:
        assert c[k].shape == ho[k].shape
        if imex:
            assert errs[k] < TOL, k
        else:
            assert not c[k].any()
    assert abs(c["time"] - o.diag()["t"]) < 1e-18
    for f in FIELDS:
        assert rel(c[f], so[f]) < TOL


@pytest.mark.parametrize("scheme", ["CNAB2", "ARS443"])
def test_snapshot_file(fd, orc, tmp_path, scheme):
    from fdns import files
    g = fd.Run(Nr, Nm, 3e5, 0.7, 1e-5, scheme)
    g.seed(SEED)
    g.step(3)
    g.write_snapshot(tmp_path / "snapshot_plot_00001")
    s = files.read_snapshot(tmp_path / "snapshot_plot_00001")
    assert (s["Ra"], s["Pr"], s["eta"], s["dt"], s["Nm"], s["Nr"]) == (3e5, 0.7, ETA, 1e-5, Nm, Nr)
    assert s["time"] == g.diag()["t"]
    assert np.array_equal(s["radius"], orc.grid(Nr, ETA)[0])
    st = g.get_state()
    for f in FIELDS:
        assert np.array_equal(s[f], st[f])
    assert ("omgFR_check" in s) == (scheme == "ARS443")


@pytest.mark.parametrize("scheme", ["CNAB2", "SBDF2", "SBDF3", "SBDF4", "ARS443", "BPR353"])
def test_restart_continues_run(fd, orc, scheme):
    """Restart from the last K states of a run (K = n_order_tscheme_imp for
    IMEX, 1 for IMEXRK), then M more steps: equal to the run that went on."""
    S, M, dt = 6, 4, 1e-5
    A = fd.Run(Nr, Nm, 1e6, 1.0, dt, scheme)
    A.seed(SEED)
    states, times = [], []
    for _ in range(S):
        A.step(1)
        g = A.get_state()
        states.append(tuple(g[f] for f in FIELDS))
        times.append(A.diag()["t"])
    A.step(M)
    ref = A.get_state()
    ni, _ = A.orders()
    K = ni if fd.SCHEMES[scheme][0] == "IMEX" else 1
    B = fd.Run(Nr, Nm, 1e6, 1.0, dt, scheme, restart=True)
    B.restart(states[-K:], times[-1])
    B.step(M)
    got = B.get_state()
    assert B.diag()["t"] == A.diag()["t"]
    e_run = max(rel(got[f], ref[f]) for f in FIELDS)
    o = orc.Run(Nr, Nm, 1e6, 1.0, dt, scheme, restart=True)
    o.restart(states[-K:], times[-1])
    o.step(M)
    e_orc = max(rel(got[f], o.get_state()[f]) for f in FIELDS)
    print(f"{scheme} K={K}: restart vs uninterrupted {e_run:.2e}, vs oracle restart {e_orc:.2e}", flush=True)
    assert e_run < 1e-11
    assert e_orc < TOL


def test_restart_single_checkpoint_imex_vs_oracle(fd, orc):
    """IMEX without l_imexrk_started: one checkpoint, zero histories."""
    dt = 1e-5
    A = fd.Run(Nr, Nm, 1e6, 1.0, dt, "SBDF3")
    A.seed(SEED)
    A.step(3)
    g = A.get_state()
    st = tuple(g[f] for f in FIELDS)
    t = A.diag()["t"]
    B = fd.Run(Nr, Nm, 1e6, 1.0, dt, "SBDF3", restart=True)
    B.restart([st], t)
    B.step(4)
    o = orc.Run(Nr, Nm, 1e6, 1.0, dt, "SBDF3", restart=True)
    o.restart([st], t)
    o.step(4)
    got, want = B.get_state(), o.get_state()
    assert max(rel(got[f], want[f]) for f in FIELDS) < TOL
    with pytest.raises(fd.FdnsError):
        B.restart([st, st], t)           # neither 1 nor n_order_tscheme_imp = 3


def test_restart_needs_restart_run(fd):
    g = fd.Run(Nr, Nm, 1e6, 1.0, 1e-5, "CNAB2")
    g.seed(SEED)
    st = tuple(g.get_state()[f] for f in FIELDS)
    with pytest.raises(fd.FdnsError):
        g.restart([st, st], 0.0)


# ------------------------------------------------------------------ driver
def _driver_input(tmp_path, name, **kw):
    from test_driver import write_input
    inp = write_input(tmp_path / name, Nr=Nr, Nm=Nm, seed=SEED, **kw)
    return inp


def _set(inp, **subs):
    text = inp.read_text()
    for k, v in subs.items():
        text = text.replace(k, v)
    inp.write_text(text)


@pytest.mark.parametrize("stype,imp,exp,scheme", [("IMEX", "BDF3", "AB3", "SBDF3"),
                                                  ("IMEXRK", "ARS443", "expARS443", "ARS443")])
def test_driver_files_equal_library_files(fd, tmp_path, stype, imp, exp, scheme):
    """The driver's Fortran-written checkpoint and snapshot files, byte for
    byte, against the library's writers on the same run."""
    from test_driver import driver
    inp = _driver_input(tmp_path, "input.txt", steps=6, stype=stype, imp=imp, exp=exp, every=3)
    _set(inp, **{"n_checkpoint=10000": "n_checkpoint=3", "n_snapshot=100000": "n_snapshot=3",
                 "n_restart_point=1": "n_restart_point=0"})
    p = subprocess.run([driver(), str(inp)], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stderr
    g = fd.Run(Nr, Nm, 1e5, 1.0, 1e-5, scheme)
    g.seed(SEED, amp=0.001)
    for k in (1, 2):
        g.step(3)
        g.write_checkpoint(tmp_path / f"lib_checkpoint_{k}")
        g.write_snapshot(tmp_path / f"lib_snapshot_{k}")
        assert (tmp_path / f"lib_checkpoint_{k}").read_bytes() == (tmp_path / f"checkpoint_{k:05d}").read_bytes()
        assert (tmp_path / f"lib_snapshot_{k}").read_bytes() == (tmp_path / f"snapshot_plot_{k:05d}").read_bytes()


def test_driver_restart_continues_series(tmp_path):
    """Driver run of 8 steps with a checkpoint every step; a second run
    restarted at step 6 (BDF3 reads checkpoints 4, 5, 6) writes the same KE and
    Nusselt rows for steps 7 and 8."""
    from test_driver import driver
    a = _driver_input(tmp_path, "a.txt", steps=8, stype="IMEX", imp="BDF3", exp="AB3", every=1, tag="a")
    _set(a, **{"n_checkpoint=10000": "n_checkpoint=1", "n_restart_point=1": "n_restart_point=0"})
    p = subprocess.run([driver(), str(a)], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stderr
    b = _driver_input(tmp_path, "b.txt", steps=2, stype="IMEX", imp="BDF3", exp="AB3", every=1, tag="b")
    _set(b, **{"n_checkpoint=10000": "n_checkpoint=1", "l_restart=.FALSE.": "l_restart=.TRUE.",
               "l_imexrk_started = .FALSE.": "l_imexrk_started = .TRUE.", "n_restart=0": "n_restart=6",
               "n_restart_point=1": "n_restart_point=6"})
    p = subprocess.run([driver(), str(b)], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stderr
    ka = np.loadtxt(tmp_path / "KE_a", ndmin=2)
    kb = np.loadtxt(tmp_path / "KE_b", ndmin=2)
    na = np.loadtxt(tmp_path / "Nusselt_a", ndmin=2)
    nb = np.loadtxt(tmp_path / "Nusselt_b", ndmin=2)
    assert ka.shape == (8, 6) and kb.shape == (2, 6)
    # the files carry 13 significant digits (es20.12)
    assert np.allclose(kb[:, :5], ka[6:, :5], rtol=1e-11, atol=0)
    # the mean-flow energy is O(amp^4) this early, on the scale of the total
    assert np.allclose(kb[:, 5], ka[6:, 5], rtol=0, atol=1e-11 * ka[:, 2].max())
    assert np.allclose(nb, na[6:], rtol=1e-11, atol=0)


def test_refine_state_vs_oracle(fd, orc):
    """fdns_refine_state against the oracle's restatement of init.F90:722-861
    (the C-ABI's host arithmetic; run here with the other restart tests)."""
    rng = np.random.default_rng(12)
    st = tuple(np.asfortranarray(rng.standard_normal((9, 17)) + 1j * rng.standard_normal((9, 17))) for _ in range(4))
    for Nm_new, Nr_new in ((8, 17), (16, 17), (8, 33), (16, 33)):
        a = fd.refine_state(st, Nm_new, Nr_new)
        b = orc.refine(st, Nm_new, Nr_new)
        for x, y in zip(a, b):
            assert x.shape == (Nm_new + 1, Nr_new)
            assert np.abs(x - y).max() <= 1e-13 * np.abs(y).max()
    with pytest.raises(fd.FdnsError):
        fd.refine_state(st, 4, 17)


def test_driver_restart_on_finer_grid(fd, orc, tmp_path):
    """A 17 x 8 checkpoint restarted by the driver on 33 x 16: the KE and
    Nusselt rows equal those of the library run restarted from the oracle's
    refinement of that checkpoint."""
    from test_driver import driver
    from fdns import files
    a = _driver_input(tmp_path, "a.txt", steps=4, stype="IMEXRK", imp="ARS443", exp="expARS443", every=4, tag="a")
    text = a.read_text().replace("Nr_max = 33", "Nr_max = 17").replace("Nm_max = 16", "Nm_max = 8")
    a.write_text(text)
    _set(a, **{"n_checkpoint=10000": "n_checkpoint=4", "n_restart_point=1": "n_restart_point=0"})
    p = subprocess.run([driver(), str(a)], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stderr
    c = files.read_checkpoint(tmp_path / "checkpoint_00001")
    assert (c["Nm"], c["Nr"]) == (8, 17)
    b = _driver_input(tmp_path, "b.txt", steps=3, stype="IMEXRK", imp="ARS443", exp="expARS443", every=1, tag="b")
    _set(b, **{"l_restart=.FALSE.": "l_restart=.TRUE.", "n_restart=0": "n_restart=4",
               "n_restart_point=1": "n_restart_point=1"})
    p = subprocess.run([driver(), str(b)], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stderr
    kb = np.loadtxt(tmp_path / "KE_b", ndmin=2)
    nb = np.loadtxt(tmp_path / "Nusselt_b", ndmin=2)
    st = orc.refine(tuple(c[f] for f in FIELDS), Nm, Nr)
    r = fd.Run(Nr, Nm, 1e5, 1.0, 1e-5, "ARS443", restart=True)
    r.restart([st], c["time"])
    rows = []
    for _ in range(3):
        r.step(1)
        d = r.diag()
        rows.append((d["t"], d["KE"], d["Nu_t"], d["Nu_b"]))
    rows = np.array(rows)
    assert kb.shape == (3, 6)
    # the files carry 13 significant digits (es20.12)
    assert np.allclose(kb[:, 1:3], rows[:, :2], rtol=1e-11, atol=0)
    assert np.allclose(nb, rows[:, [0, 2, 3]], rtol=1e-11, atol=0)
