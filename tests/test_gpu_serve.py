"""The serving pool at the served stream's job sizes, and its multi-step calls.

* fdns_pool_run (several batched steps per call, one captured graph per
  composition, one host synchronisation per call) gives every job the same
  result, bit for bit, as stepping the pool one step per call and as the job
  run alone.
* A Poisson-ordered stream at the sizes of the served workload -- Nr in
  {33, 65, 129, 257} (LU solves at 257), Nm up to 768, CNAB2 / SBDF3 / ARS443 /
  BPR353, Ra 1e5..1e8, Pr in {0.7, 1, 7} -- with jobs arriving while others
  run: each result equals the job alone and the oracle within 1e-10.
"""
from __future__ import annotations

import numpy as np
import pytest

from conftest import rel
from test_gpu_batch import MIX, pool_alone

pytestmark = pytest.mark.gpu


def _submit(P, j):
    return P.submit(j["Nr"], j["Nm"], j["Ra"], j["Pr"], j["dt"], j["scheme"], j["steps"], j["seed"],
                    **{x: j[x] for x in ("restart", "mBC") if x in j})


def _serve(fd, jobs, arrivals, budget, max_steps):
    """arrivals: call index -> jobs submitted before that call."""
    P = fd.Pool(budget)
    ids = {}
    calls = 0
    while True:
        for k in arrivals.get(calls, []):
            ids[k] = _submit(P, jobs[k])
        if max_steps == 1:
            P.step()
        else:
            P.run(max_steps)
        calls += 1
        if calls > max(arrivals) and all(P.poll(i)[0] == 2 for i in ids.values()):
            break
        assert calls < 2000
    out = {}
    for k in range(len(jobs)):
        status, done = P.poll(ids[k])
        assert status == 2 and done == jobs[k]["steps"]
        out[k] = P.result(ids[k])
    P.close()
    return out, calls


def test_pool_run_equals_pool_step(fd):
    jobs = []
    for k, m in enumerate(MIX):
        j = dict(m)
        j["steps"] = 7 + 5 * k
        jobs.append(j)
    budget = 1 << 36
    arrivals = {0: [0, 1, 2, 3], 1: [4, 5], 3: [6, 7], 4: [8, 9]}
    one, calls1 = _serve(fd, jobs, arrivals, budget, 1)
    many, calls8 = _serve(fd, jobs, arrivals, budget, 8)
    assert calls8 < calls1
    for k, j in enumerate(jobs):
        for f in ("tFR", "omgFR", "urFR", "upFR"):
            assert np.array_equal(one[k][f], many[k][f]), (k, f)
    alone = pool_alone(fd, jobs[3], budget)
    assert np.array_equal(alone["tFR"], many[3]["tFR"])


def c3_stream():
    """Ten jobs at the served workload's sizes, in the order a Poisson stream
    brings them, a few steps each."""
    sizes = [(257, 768, "ARS443"), (33, 64, "CNAB2"), (129, 512, "SBDF3"), (65, 256, "BPR353"),
             (257, 128, "CNAB2"), (33, 768, "ARS443"), (129, 64, "BPR353"), (65, 512, "CNAB2"),
             (97, 256, "SBDF3"), (257, 256, "BPR353")]
    rng = np.random.default_rng(33)
    jobs = []
    for k, (Nr, Nm, sch) in enumerate(sizes):
        jobs.append(dict(Nr=Nr, Nm=Nm, Ra=float(10 ** rng.uniform(5, 8)), Pr=float(rng.choice([0.7, 1.0, 7.0])),
                         dt=1e-7, scheme=sch, steps=int(rng.integers(3, 7)), seed=4000 + k,
                         restart=(sch == "SBDF3")))
    return jobs


def test_c3_stream_vs_alone_and_oracle(fd, orc):
    jobs = c3_stream()
    # a budget that holds a few jobs at a time, so later arrivals wait
    budget = 6 << 30
    arrivals = {0: [0, 1, 2], 1: [3, 4], 2: [5], 3: [6, 7], 5: [8, 9]}
    mixed, _ = _serve(fd, jobs, arrivals, budget, 4)
    for k, j in enumerate(jobs):
        alone = pool_alone(fd, j, budget)
        for f in ("tFR", "omgFR", "urFR", "upFR"):
            assert np.array_equal(mixed[k][f], alone[f]), (k, f)
        o = orc.Run(j["Nr"], j["Nm"], j["Ra"], j["Pr"], j["dt"], j["scheme"], restart=j["restart"])
        o.set_state(*orc.seed_state(j["Nr"], j["Nm"], 0.35, j["seed"], amp=1e-3))
        o.step(j["steps"])
        s = o.get_state()
        for f in ("tFR", "omgFR", "urFR", "upFR"):
            assert rel(mixed[k][f], s[f]) < 1e-10, (k, j["Nr"], j["Nm"], j["scheme"], f, rel(mixed[k][f], s[f]))
