"""Long horizons on the GPU against the oracle.

* Three members of the 64-run mixed ensemble (CNAB2, restarted SBDF3 and
  ARS443 at Nr = 97, Nm = 256, dt = 1e-6) advanced for the workload's full
  1000 steps by one batched ensemble -- inverse-mode solves, replays of the
  captured step -- against the oracle (dgetrs-order solves): the fields and
  the (t, KE, Nu_t, Nu_b) series every 50 steps within 1e-10.
* The solver crossover: at Nr = 257 over 2000 steps the LU-order solves stay
  within 1e-10 of the oracle, and the inverse-mode error is printed beside
  them (the Nusselt number's wall derivative amplifies the inverse product's
  rounding by ~Nr^2; the automatic choice takes LU order above Nr = 129).
"""
from __future__ import annotations

import os
import sys

import numpy as np
import pytest

from conftest import REPO, rel

pytestmark = pytest.mark.gpu

TOL = 1e-10
ETA = 0.35
EVERY = 50
FIELDS = ("tFR", "omgFR", "urFR", "upFR", "psi")
THREADS = min(16, os.cpu_count() or 1)


def series_err(a, b):
    a, b = np.asarray(a), np.asarray(b)
    assert a.shape == b.shape, (a.shape, b.shape)
    assert np.allclose(a[:, 0], b[:, 0], rtol=1e-14, atol=0)
    return max(float(np.max(np.abs(a[:, c] - b[:, c]) / np.maximum(np.abs(b[:, c]), 1e-300))) for c in (1, 2, 3))


def oracle_trajectory(orc, m, steps, every):
    o = orc.Run(m["Nr"], m["Nm"], m["Ra"], m["Pr"], m["dt"], m["scheme"], restart=m.get("restart", False),
                nthreads=THREADS)
    o.set_state(*orc.seed_state(m["Nr"], m["Nm"], ETA, m["seed"]))
    rows = []
    for _ in range(steps // every):
        o.step(every)
        d = o.diag()
        rows.append((d["t"], d["KE"], d["Nu_t"], d["Nu_b"]))
    return o.get_state(), np.array(rows)


def c2_members():
    sys.path.insert(0, REPO)
    from bench import ensemble_members
    return ensemble_members()[:3]


@pytest.fixture(scope="module")
def c2_ensemble(fd):
    """The three members stepped 1000 times as one ensemble."""
    runs = []
    for m in c2_members():
        r = fd.Run(m["Nr"], m["Nm"], m["Ra"], m["Pr"], m["dt"], m["scheme"], restart=m["restart"], n_KE=EVERY)
        r.seed(m["seed"])
        runs.append(r)
    ens = fd.Ensemble(runs)
    ens.step(1000, graph=True)
    out = []
    for r in runs:
        s = r.get_state()
        s["series"] = r.series()
        out.append(s)
    ens.close()
    for r in runs:
        r.close()
    return out


@pytest.mark.parametrize("k", [0, 1, 2])
def test_c2_member_1000_steps(c2_ensemble, orc, k):
    m = c2_members()[k]
    want, series = oracle_trajectory(orc, m, 1000, EVERY)
    got = c2_ensemble[k]
    errs = {f: rel(got[f], want[f]) for f in FIELDS}
    es = series_err(got["series"][:, :4], series)
    print(f"c2 member {k} ({m['scheme']}, Ra={m['Ra']:.3g}, Pr={m['Pr']}) 1000 steps:",
          {f: f"{v:.2e}" for f, v in errs.items()}, f"series {es:.2e}", flush=True)
    assert max(errs.values()) <= TOL
    assert es <= TOL


def test_solver_crossover_nr257(fd, orc):
    m = dict(Nr=257, Nm=16, Ra=1e8, Pr=1.0, dt=1e-7, scheme="CNAB2", seed=1257)
    steps = 2000
    want, series = oracle_trajectory(orc, m, steps, 200)
    res = {}
    for solver in ("lu", "inverse"):
        g = fd.Run(m["Nr"], m["Nm"], m["Ra"], m["Pr"], m["dt"], m["scheme"], n_KE=200, solver=solver)
        g.seed(m["seed"])
        g.step(steps)
        s = g.get_state()
        res[solver] = ({f: rel(s[f], want[f]) for f in FIELDS}, series_err(g.series()[:, :4], series))
        g.close()
        print(f"Nr=257 Nm=16 CNAB2 {steps} steps, {solver} solves:",
              {f: f"{v:.2e}" for f, v in res[solver][0].items()}, f"KE/Nu series {res[solver][1]:.2e}", flush=True)
    assert max(res["lu"][0].values()) <= TOL and res["lu"][1] <= TOL
