"""The reference's own start-up on the GPU, through the C-ABI and the Fortran
driver: init_perturbation (init.F90:277-399), sizes whose transform lengths
have a prime factor above 5 (the reference's input.txt has Nr_max = 32, a
radial length 2*32-2 = 62 = 2*31), the dt_array of a restart
(init.F90:464-612), the namelist defaults (namelists.F90:112-146) and the
totaltime exit (steptime.F90:255, 287-292).
"""
from __future__ import annotations

import os
import subprocess

import numpy as np
import pytest

from conftest import rel
from test_driver import INPUT, driver

pytestmark = pytest.mark.gpu
ETA = 0.35


@pytest.mark.parametrize("N,Nm,n_init", [(32, 32, 0), (32, 32, -1), (33, 16, 4), (20, 7, 0), (257, 768, 0),
                                         (97, 256, -1)])
def test_init_perturbation_vs_oracle(fd, orc, N, Nm, n_init):
    run = fd.Run(N, Nm, 2000.0, 1.0, 1e-5, "CNAB2")
    run.init_perturbation(n_init, 0.01)
    g = run.get_state()
    o = orc.init_perturbation(N, Nm, ETA, n_init, 0.01)
    assert rel(g["tFR"], o[0]) < 1e-14
    for k in ("omgFR", "urFR", "upFR"):
        assert not g[k].any()
    run.close()


# the reference's src/input.txt: 32 x 32, Ra 2000, Pr 1, no slip, ampT 0.01,
# lagpts 16, n_init 0, IMEXRK BHR553, variable dt (CFL 0.1, dt_coef 1,
# dt_max 1e-2) from dt = 1e-5
INPUT_TXT = dict(Nr=32, Nm=32, Ra=2000.0, Pr=1.0, dt=1e-5, scheme="BHR553", lagpts=16, vardt=True, CFL=0.1,
                 dt_coef=1.0, dt_max=1e-2)


def test_reference_input_vs_oracle(fd, orc):
    c = INPUT_TXT
    kw = dict(vardt=c["vardt"], CFL=c["CFL"], dt_coef=c["dt_coef"], dt_max=c["dt_max"], lagpts=c["lagpts"])
    g = fd.Run(c["Nr"], c["Nm"], c["Ra"], c["Pr"], c["dt"], c["scheme"], **kw)
    g.init_perturbation(0, 0.01)
    o = orc.Run(c["Nr"], c["Nm"], c["Ra"], c["Pr"], c["dt"], c["scheme"], **kw)
    o.set_state(*orc.init_perturbation(c["Nr"], c["Nm"], ETA, 0, 0.01))
    for n in range(30):
        g.step(1)
        o.step(1)
        assert abs(g.dt() - o.dt()) <= 1e-13 * o.dt(), n
    a, b = g.get_state(), o.get_state()
    assert o.dt() > 1e-5                      # the CFL limit let dt grow
    for k in ("tFR", "omgFR", "urFR", "upFR"):
        assert rel(a[k], b[k]) < 1e-10, k
    assert np.abs(b["urFR"]).max() > 1e-8     # the bubble has set the fluid moving
    g.close()


@pytest.mark.parametrize("solver", ["lu", "inverse"])
@pytest.mark.parametrize("N,Nm,scheme", [(20, 7, "CNAB2"), (32, 32, "ARS443"), (44, 14, "SBDF3"), (32, 11, "RK4")])
def test_generic_radix_sizes_vs_oracle(fd, orc, N, Nm, scheme, solver):
    """2N-2 = 38, 62, 86 (primes 19, 31, 43) and 3 Nm = 21, 42, 33 (7, 11):
    the direct-DFT pass of the transforms."""
    dt = 2e-6 if scheme == "RK4" else 1e-5
    restart = scheme == "SBDF3"
    g = fd.Run(N, Nm, 1e5, 1.0, dt, scheme, restart=restart, solver=solver)
    g.init_perturbation(3, 0.3)
    o = orc.Run(N, Nm, 1e5, 1.0, dt, scheme, restart=restart)
    o.set_state(*orc.init_perturbation(N, Nm, ETA, 3, 0.3))
    g.step(10)
    o.step(10)
    a, b = g.get_state(), o.get_state()
    for k in ("tFR", "omgFR", "urFR", "upFR"):
        assert rel(a[k], b[k]) < 1e-10, k
    g.close()


def test_restart_dt_array_vs_oracle(fd, orc):
    """A variable-dt SBDF3 restart from three checkpoint states with their
    dt_array restored, GPU against the oracle."""
    kw = dict(restart=True, vardt=True, CFL=0.5, dt_coef=1.0, dt_max=4e-4)
    o = orc.Run(17, 8, 2e4, 1.0, 1e-4, "SBDF3", **kw)
    o.set_state(*orc.seed_state(17, 8, ETA, 3, amp=0.5))
    states, dts = [], []
    for _ in range(9):
        o.step(1)
        s = o.get_state()
        states.append((s["tFR"], s["omgFR"], s["urFR"], s["upFR"]))
        dts.append(o.dt())
    t_last = o.diag()["t"]
    dta = [dts[-1], dts[-2], dts[-3]]
    assert len(set(dta)) == 3
    g = fd.Run(17, 8, 2e4, 1.0, 1e-4, "SBDF3", **kw)
    g.restart(states[-3:], t_last)
    g.restart_dt(dta)
    o2 = orc.Run(17, 8, 2e4, 1.0, 1e-4, "SBDF3", **kw)
    o2.restart(states[-3:], t_last)
    o2.restart_dt(dta)
    for n in range(4):
        g.step(1)
        o2.step(1)
        assert abs(g.dt() - o2.dt()) <= 1e-13 * o2.dt()
    assert rel(g.get_state()["tFR"], o2.get_state()["tFR"]) < 1e-12
    g.close()


# ----------------------------------------------------------------- the driver
def _input(tmp_path, **kw):
    args = dict(Nr=32, Nm=32, Ra=2000.0, Pr=1.0, mBC="NS", buo="IMP", steps=20, dt=1e-5, vardt="TRUE",
                stype="IMEXRK", imp="BHR553", exp="expBHR553", tag="BHR553", every=10, seed=-1)
    args.update(kw)
    text = INPUT.format(**args).replace("ampT = 0.001,", "ampT = 0.01,")
    if args["seed"] < 0:
        text = text[:text.index("&fdns")]
    (tmp_path / "input.txt").write_text(text)
    return tmp_path / "input.txt"


def test_driver_reference_input_fresh_start(fd, orc, tmp_path):
    """The reference's input.txt settings with no &fdns group: the driver starts
    from init_perturbation's bubble, and its KE and Nusselt rows follow the
    oracle's run of the same start."""
    inp = _input(tmp_path)
    p = subprocess.run([driver(), str(inp)], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stderr
    ke = np.loadtxt(tmp_path / "KE_BHR553", ndmin=2)
    nu = np.loadtxt(tmp_path / "Nusselt_BHR553", ndmin=2)
    assert ke.shape == (2, 6)
    c = INPUT_TXT
    start = orc.init_perturbation(c["Nr"], c["Nm"], ETA, 0, 0.01)
    # the Nusselt numbers of this start are sensitive to rounding: the
    # oracle's own move by up to ~5e-10 when its initial state changes by one
    # part in 1e16 (KE by 1e-15), so the Nusselt rows are held to 1e-10 or to
    # twice the largest such move over four nudges, whichever is larger
    starts = [start]
    for seed in range(1, 5):
        rng = np.random.default_rng(seed)
        starts.append(tuple(f * (1 + 1e-16 * rng.standard_normal(f.shape)) for f in start))
    rows = []
    for st in starts:
        o = orc.Run(c["Nr"], c["Nm"], c["Ra"], c["Pr"], c["dt"], c["scheme"], vardt=True, CFL=0.1, dt_coef=1.0,
                    dt_max=1e-2, lagpts=16)
        o.set_state(*st)
        r = []
        for _ in range(2):
            o.step(10)
            d = o.diag()
            r.append((o.dt(), d["t"], d["KE"], d["Nu_t"], d["Nu_b"]))
        rows.append(np.array(r))
    want = rows[0]
    sens = max((np.abs(o - want) / np.abs(want))[:, 3:].max() for o in rows[1:])
    tol = max(1e-10, 2 * sens)
    for row in range(2):
        assert abs(ke[row, 0] - want[row, 0]) <= 1e-11 * want[row, 0]
        assert abs(nu[row, 0] - want[row, 1]) <= 1e-11 * want[row, 1]
        assert abs(ke[row, 2] - want[row, 2]) <= 1e-10 * want[row, 2]
        for col, k in ((1, 3), (2, 4)):
            assert abs(nu[row, col] - want[row, k]) <= tol * abs(want[row, k]), (row, col, tol)


def test_driver_output_defaults(fd, tmp_path):
    """An &output group that gives only tag and n_KE: the defaults of
    n_checkpoint, n_snapshot and n_KEspec (500000, 300000, 100000) write no
    checkpoint, snapshot or spectra in 1000 steps."""
    inp = _input(tmp_path, Nr=17, Nm=8, Ra=1e4, vardt="FALSE", stype="IMEX", imp="CN", exp="AB2", tag="d",
                 every=500, steps=1000, dt=1e-6, seed=5)
    text = inp.read_text()
    for line in ("n_checkpoint=10000,\n", "n_snapshot=100000,\n", "n_KEspec=100000\n"):
        assert line in text
        text = text.replace(line, "")
    inp.write_text(text)
    p = subprocess.run([driver(), str(inp)], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stderr
    assert np.loadtxt(tmp_path / "KE_d", ndmin=2).shape[0] == 2
    assert not [f for f in os.listdir(tmp_path) if f.startswith(("checkpoint_", "snapshot_plot_"))]
    assert os.path.getsize(tmp_path / "KEspec_d") == 0


@pytest.mark.parametrize("stype,imp,exp,steps_run,rows", [("IMEX", "CN", "AB2", 11, 2),
                                                          ("RK", "NONE", "RK4", 40, 10)])
def test_driver_totaltime(fd, tmp_path, stype, imp, exp, steps_run, rows):
    """totaltime = 10.5 dt: the IMEX loop ends after step 11, the first whose
    time reaches it, writing no row for it; the RK loop has no such exit."""
    inp = _input(tmp_path, Nr=33, Nm=8, Ra=1e4, vardt="FALSE", stype=stype, imp=imp, exp=exp, tag="tt",
                 every=4, steps=40, dt=1e-5 if stype != "RK" else 1e-6, seed=5)
    total = 10.5 * (1e-5 if stype != "RK" else 1e-6)
    inp.write_text(inp.read_text().replace("totaltime=100.0", f"totaltime={total!r}"))
    p = subprocess.run([driver(), str(inp)], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stderr
    assert f"steps: {steps_run} " in p.stdout
    assert np.loadtxt(tmp_path / "KE_tt", ndmin=2).shape[0] == rows
