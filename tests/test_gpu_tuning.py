"""Launch choices (fdns_tuning_get / fdns_tuning_set) and the autotuner
(fdns_ens_autotune) on the GPU.

* Every launch choice -- nl_tpb 128 / 256, tile_waves 4 / 8, tile_stage 0 / 1,
  tile_pf 0 / 1, solve_chunk largest / 16 / 8 -- gives bit for bit the states
  of the default
  choice, for IMEX and IMEXRK groups, inverse and LU solves, graph replays
  and eager steps.  At Nr = 257 the three-block column-tile kernel does not
  fit its staged blocks in LDS and keeps direct loads while the others stage:
  both paths run in one step there.
* A choice changed between steps reaches an ensemble whose step graph was
  captured before (the graph is captured again).
* The autotuner returns a valid choice, stores it for the shape, and leaves
  the ensemble's members untouched.
"""
from __future__ import annotations

import itertools

import numpy as np
import pytest

from test_gpu_batch import make, states_equal

pytestmark = pytest.mark.gpu

DEFAULT = dict(nl_tpb=256, tile_waves=8, tile_stage=0, solve_chunk=0)
MIXED = dict(nl_tpb=128, tile_waves=[4, 8, 4, 8, 4, 8], tile_stage=[1, 0, 0, 1, 1, 0], solve_chunk=16,
             tile_pf=[1, 0, 1, 1, 0, 1], lane_split=1)

# members of one shape in three groups: fresh IMEX, restarted IMEX, IMEXRK
def members(Nr, Nm, n):
    sch = [("CNAB2", False), ("SBDF3", True), ("ARS443", False), ("BPR353", False)]
    out = []
    for k in range(n):
        s, rst = sch[k % len(sch)]
        out.append(dict(Nr=Nr, Nm=Nm, Ra=1e5 * (1 + k % 3), Pr=(0.7, 1.0, 7.0)[k % 3], dt=1e-6, scheme=s,
                        seed=300 + k, restart=rst))
    return out


def run_ens(fd, ms, steps, graph=True, tuning=None, change=None):
    """States after `steps` batched steps; `change` = (after_step, tuning)
    switches the choice between steps."""
    Nr, Nm = ms[0]["Nr"], ms[0]["Nm"]
    prev = fd.get_tuning(Nr, Nm)
    fd.set_tuning(Nr, Nm, **(tuning or DEFAULT))
    runs = [make(fd, m) for m in ms]
    ens = fd.Ensemble(runs)
    try:
        if change is None:
            ens.step(steps, graph=graph)
        else:
            k, t2 = change
            ens.step(k, graph=graph)
            fd.set_tuning(Nr, Nm, **t2)
            ens.step(steps - k, graph=graph)
        return [r.get_state() for r in runs]
    finally:
        ens.close()
        for r in runs:
            r.close()
        fd.set_tuning(Nr, Nm, **prev)


CHOICES = [dict(nl_tpb=a, tile_waves=b, tile_stage=c, solve_chunk=d, tile_pf=e)
           for a, b, c, d, e in itertools.product((128, 256), (4, 8), (0, 1), (0, 16, 8), (0, 1))]


@pytest.mark.parametrize("shape", [(65, 128, 20), (97, 256, 9), (257, 48, 6)])
def test_every_choice_same_bits(fd, shape):
    Nr, Nm, n = shape
    ms = members(Nr, Nm, n)
    ref = run_ens(fd, ms, 4)
    # every family's options at least once, in combinations
    picks = (CHOICES if Nr == 65 else [CHOICES[i] for i in (1, 10, 21, 30, 39, 44, 47)]) + [MIXED]
    for t in picks:
        got = run_ens(fd, ms, 4, tuning=t)
        for k, (a, b) in enumerate(zip(got, ref)):
            assert states_equal(a, b), (t, k, ms[k]["scheme"])


def test_choice_change_reaches_captured_graph(fd):
    ms = members(97, 64, 8)
    ref = run_ens(fd, ms, 6, graph=True)
    staged = dict(nl_tpb=128, tile_waves=4, tile_stage=1, solve_chunk=8, tile_pf=1, lane_split=1)
    got = run_ens(fd, ms, 6, graph=True, change=(3, staged))
    for a, b in zip(got, ref):
        assert states_equal(a, b)
    # after the change the ensemble runs the new kernels: the same steps
    # eagerly with the new choice from the start give the same bits too
    eager = run_ens(fd, ms, 6, graph=False, tuning=staged)
    for a, b in zip(eager, ref):
        assert states_equal(a, b)


def test_solo_run_follows_choice(fd):
    m = members(65, 64, 1)[0]
    prev = fd.get_tuning(65, 64)
    fd.set_tuning(65, 64, **DEFAULT)
    a = make(fd, m)
    a.step(3)
    fd.set_tuning(65, 64, nl_tpb=128, tile_waves=4, tile_stage=1, solve_chunk=16, tile_pf=1)
    a.step(3)
    fd.set_tuning(65, 64, **prev)
    b = make(fd, m)
    b.step(6)
    assert states_equal(a.get_state(), b.get_state())
    a.close()
    b.close()


def test_autotune(fd):
    ms = members(65, 128, 12)
    prev = fd.get_tuning(65, 128)
    ref = run_ens(fd, ms, 5)
    runs = [make(fd, m) for m in ms]
    ens = fd.Ensemble(runs)
    ens.step(2)
    chosen, report = ens.autotune(steps=2)
    assert chosen["nl_tpb"] in (128, 256) and set(chosen["tile_waves"]) <= {4, 8}
    assert set(chosen["tile_stage"]) <= {0, 1} and set(chosen["tile_pf"]) <= {0, 1} and chosen["lane_split"] in (0, 1)
    assert chosen["solve_chunk"] in (0, 16, 8)
    assert fd.get_tuning(65, 128) == chosen
    # every tile kernel the ensemble launches was timed with all eight options
    for k in fd.TILE_KERNELS:
        assert sum(1 for key in report if key.startswith(k + "/")) == 8, k
    assert all(v > 0 for v in report.values())
    ens.step(3)
    got = [r.get_state() for r in runs]
    ens.close()
    for r in runs:
        r.close()
    fd.set_tuning(65, 128, **prev)
    for a, b in zip(got, ref):
        assert states_equal(a, b)
