"""Batching and serving on the GPU: a run's result does not depend on what it
is batched with.

* Ensemble: every member of a mixed ensemble (several shapes and schemes in
  one batched step) equals the same run stepped alone, bit for bit.
* hipGraph replay: a captured batched step replayed many times, with eager
  work and synchronisations between replays, equals eager launches bit for bit.
* Serving pool: jobs arriving while others run are admitted as budget frees
  up; each job's result equals the job run alone, bit for bit ("alone and in
  the mix").
"""
from __future__ import annotations

import numpy as np
import pytest

from conftest import rel

pytestmark = pytest.mark.gpu

MIX = [
    dict(Nr=33, Nm=16, Ra=1e5, Pr=1.0, dt=1e-5, scheme="CNAB2", seed=1),
    dict(Nr=33, Nm=16, Ra=3e5, Pr=7.0, dt=1e-5, scheme="SBDF3", seed=2, restart=True),
    dict(Nr=33, Nm=16, Ra=2e5, Pr=0.7, dt=1e-5, scheme="ARS443", seed=3),
    dict(Nr=65, Nm=32, Ra=1e6, Pr=1.0, dt=1e-6, scheme="BPR353", seed=4),
    dict(Nr=65, Nm=32, Ra=1e6, Pr=1.0, dt=1e-6, scheme="CNAB2", seed=5),
    dict(Nr=33, Nm=16, Ra=1e5, Pr=1.0, dt=1e-5, scheme="CNAB2", seed=6),
    dict(Nr=17, Nm=64, Ra=1e5, Pr=1.0, dt=1e-5, scheme="SBDF2", seed=7, restart=True),
    dict(Nr=33, Nm=16, Ra=2e5, Pr=0.7, dt=1e-5, scheme="ARS443", seed=8, mBC="SF"),
    dict(Nr=33, Nm=16, Ra=2e5, Pr=1.0, dt=1e-5, scheme="BHR553", seed=9),
    dict(Nr=33, Nm=16, Ra=1e5, Pr=1.0, dt=1e-5, scheme="KC343", seed=10, mBC="SF"),
]


def make(fd, m):
    kw = {k: m[k] for k in ("restart", "mBC", "buo", "n_KE") if k in m}
    r = fd.Run(m["Nr"], m["Nm"], m["Ra"], m["Pr"], m["dt"], m["scheme"], **kw)
    r.seed(m["seed"])
    return r


def states_equal(a, b):
    return all(np.array_equal(a[k], b[k]) for k in ("tFR", "omgFR", "urFR", "upFR", "psi"))


@pytest.mark.parametrize("graph", [False, True])
def test_ensemble_member_equals_solo(fd, graph):
    steps = 12
    runs = [make(fd, m) for m in MIX]
    ens = fd.Ensemble(runs)
    ens.step(steps, graph=graph)
    batched = [r.get_state() for r in runs]
    ens.close()
    for m, b in zip(MIX, batched):
        solo = make(fd, m)
        solo.step(steps)
        s = solo.get_state()
        solo.close()
        assert states_equal(b, s), (m, rel(b["tFR"], s["tFR"]))
    for r in runs:
        r.close()


def test_graph_replay_equals_eager(fd):
    """Replays of the captured step, each followed by eager work and a host
    synchronisation, against the same steps launched eagerly."""
    a = [make(fd, m) for m in MIX[:5]]
    b = [make(fd, m) for m in MIX[:5]]
    ea, eb = fd.Ensemble(a), fd.Ensemble(b)
    ea.step(2, graph=False)
    eb.step(2, graph=False)
    other = make(fd, MIX[5])
    for k in range(6):
        ea.step(1, graph=True)          # replay of the captured step
        eb.step(1, graph=False)
        other.step(1)                   # unrelated eager work between replays
        for x, y in zip(a, b):
            assert states_equal(x.get_state(), y.get_state()), k
    for e in (ea, eb):
        e.close()
    for r in a + b + [other]:
        r.close()


def test_graph_replays_with_records_equal_eager(fd):
    """Multi-step calls that replay the captured step between recorded steps
    (n_KE = 5, launched eagerly), with eager work between the calls, against
    the same steps launched eagerly."""
    for nKE in (0, 5):
        a = [make(fd, dict(m, n_KE=nKE)) for m in MIX[:6]]
        b = [make(fd, dict(m, n_KE=nKE)) for m in MIX[:6]]
        ea, eb = fd.Ensemble(a), fd.Ensemble(b)
        ea.step(1, graph=True)
        eb.step(1, graph=False)
        other = make(fd, MIX[6])
        for k in range(3):
            ea.step(9, graph=True)
            eb.step(9, graph=False)
            other.step(2)
            for x, y in zip(a, b):
                assert states_equal(x.get_state(), y.get_state()), (nKE, k)
        for e in (ea, eb):
            e.close()
        for r in a + b + [other]:
            r.close()


# the m = 0 temperature solves: in the first blocks of the temperature-solve
# launch (Nr_max <= 100; one 64-row slot at Nr 33, two at 65 and 97), on the
# side stream (Nr_max 129), or as their own kernel ahead of it (the per-kernel
# timed steps)
M0_MIX = [
    dict(Nr=33, Nm=32, Ra=1e5, Pr=1.0, dt=1e-6, scheme="CNAB2", seed=21),
    dict(Nr=65, Nm=64, Ra=1e6, Pr=0.7, dt=1e-6, scheme="ARS443", seed=22),
    dict(Nr=97, Nm=64, Ra=1e6, Pr=7.0, dt=1e-6, scheme="SBDF3", seed=23, restart=True),
    dict(Nr=97, Nm=64, Ra=3e6, Pr=1.0, dt=1e-6, scheme="BPR353", seed=24),
    dict(Nr=129, Nm=32, Ra=1e6, Pr=1.0, dt=1e-6, scheme="ARS443", seed=25),
]


def test_m0_solve_placements_same_bits(fd):
    ref = []
    for m in M0_MIX:
        r = make(fd, m)
        e = fd.Ensemble([r])
        e.step_timed(6)                 # k_m0, then k_solve
        ref.append(r.get_state())
        e.close()
        r.close()
    for graph in (False, True):
        runs = [make(fd, m) for m in M0_MIX]
        ens = fd.Ensemble(runs)
        ens.step(6, graph=graph)
        for m, r, s in zip(M0_MIX, runs, ref):
            assert states_equal(r.get_state(), s), (graph, m)
        ens.close()
        for r in runs:
            r.close()


def test_ensemble_rejects_members_at_different_steps(fd):
    r1, r2 = make(fd, MIX[0]), make(fd, MIX[5])
    r1.step(1)
    with pytest.raises(fd.FdnsError):
        fd.Ensemble([r1, r2])
    r1.close()
    r2.close()


def test_c2_ensemble_equals_solo(fd):
    """All 64 members of the benchmark ensemble (Nr=97, Nm=256), 6 batched
    steps (graph replay after the first), against each run alone."""
    import sys
    from conftest import REPO
    sys.path.insert(0, REPO)
    from bench import ensemble_members, make_run
    members = ensemble_members()
    runs = [make_run(m) for m in members]
    ens = fd.Ensemble(runs)
    ens.step(6, graph=True)
    batched = [r.get_state()["tFR"] for r in runs]
    ens.close()
    for r in runs:
        r.close()
    for k, m in enumerate(members):
        solo = make_run(m)
        solo.step(6)
        t = solo.get_state()["tFR"]
        solo.close()
        assert np.array_equal(batched[k], t), k


def pool_alone(fd, job, budget):
    P = fd.Pool(budget)
    jid = P.submit(job["Nr"], job["Nm"], job["Ra"], job["Pr"], job["dt"], job["scheme"], job["steps"], job["seed"],
                   **{k: job[k] for k in ("restart", "mBC") if k in job})
    while P.poll(jid)[0] != 2:
        P.step()
    out = P.result(jid)
    P.close()
    return out


def test_pool_alone_equals_in_mix(fd):
    jobs = []
    for k, m in enumerate(MIX):
        j = dict(m)
        j["steps"] = 5 + 3 * k
        jobs.append(j)
    # a budget that holds the largest job but not the largest beside a small
    # one (operator sets shared by resident jobs are charged once), so later
    # arrivals wait for earlier jobs to retire
    sizes = [fd.footprint(j["Nr"], j["Nm"], j["scheme"]) for j in jobs]
    budget = max(sizes) + min(sizes) // 2
    P = fd.Pool(budget)
    ids = {}
    arrivals = {0: [0, 1, 2], 2: [3, 4], 5: [5], 6: [6, 7], 7: [8, 9]}   # pool step -> jobs submitted
    seen_waiting = False
    t = 0
    while True:
        for k in arrivals.get(t, []):
            j = jobs[k]
            ids[k] = P.submit(j["Nr"], j["Nm"], j["Ra"], j["Pr"], j["dt"], j["scheme"], j["steps"], j["seed"],
                              **{x: j[x] for x in ("restart", "mBC") if x in j})
        active, waiting = P.step()
        seen_waiting |= waiting > 0
        t += 1
        if t > max(arrivals) and all(P.poll(i)[0] == 2 for i in ids.values()):
            break
        assert t < 500
    assert seen_waiting             # admission by footprint actually held jobs back
    for k, j in enumerate(jobs):
        status, done = P.poll(ids[k])
        assert status == 2 and done == j["steps"]
        mixed = P.result(ids[k])
        alone = pool_alone(fd, j, budget)
        for f in ("tFR", "omgFR", "urFR", "upFR"):
            assert np.array_equal(mixed[f], alone[f]), (k, f)
        # and the pool's result equals a plain single run
        solo = make(fd, j)
        solo.step(j["steps"])
        s = solo.get_state()
        solo.close()
        assert np.array_equal(mixed["tFR"], s["tFR"]), k
    P.close()


def test_pool_rejects_oversized_job(fd):
    P = fd.Pool(1 << 20)
    with pytest.raises(fd.FdnsError):
        P.submit(65, 128, 1e6, 1.0, 1e-6, "CNAB2", 10, 1)
    P.close()
