"""Measure launch choices for a list of (Nr_max, Nm_max) shapes with the
library's autotuner (fdns_ens_autotune) and print them as entries of the
library's stored table (k_tuned in csrc/fdns_host.cpp).

For every shape a small mixed ensemble (CNAB2, restarted SBDF3, ARS443,
BPR353 by default: the schemes of the served stream) is stepped twice, the
autotuner times every option on clones of it, and the ensemble's time per
step (graph replays) is then measured with the plain choice and with the
chosen one, so each line states its own gain.  Results are bit for bit the
same under every choice; only the time changes.

    python3 falcon-dns_amd/tune_shapes.py                     # the served shapes
    python3 falcon-dns_amd/tune_shapes.py --shapes 97x256 --runs 64
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import fdns  # noqa: E402

SERVED = [(nr, nm) for nr in (33, 65, 129, 257) for nm in (64, 128, 256, 512, 768)]
SCHEMES = [("CNAB2", False), ("SBDF3", True), ("ARS443", False), ("BPR353", False)]


def members(Nr, Nm, n, dt):
    out = []
    for k in range(n):
        s, rst = SCHEMES[k % len(SCHEMES)]
        out.append(fdns.Run(Nr, Nm, 10.0 ** (5 + k % 3), (0.7, 1.0, 7.0)[k % 3], dt, s, restart=rst))
        out[-1].seed(9000 + k)
    return out


def ms_per_step(ens, steps):
    """Graph-replay time per step over at least `steps` steps and about
    0.2 s of stepping."""
    ens.step(2, graph=True)   # capture (and re-capture after a change) untimed
    t0 = time.perf_counter()
    ens.step(5, graph=True)
    est = (time.perf_counter() - t0) / 5
    n = max(steps, min(4000, int(0.2 / max(est, 1e-6))))
    t0 = time.perf_counter()
    ens.step(n, graph=True)
    return 1e3 * (time.perf_counter() - t0) / n


def c_entry(Nr, Nm, t):
    arr = lambda k: "{" + ", ".join(str(v) for v in t[k]) + "}"
    return (f"    {{{Nr}, {Nm}, {{{t['nl_tpb']}, {arr('tile_waves')}, {arr('tile_stage')}, {arr('tile_pf')}, "
            f"{t['solve_chunk']}, {t['lane_split']}}}}},")


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--shapes", default=None, help="comma-separated NrxNm list (default: the served shapes)")
    ap.add_argument("--runs", type=int, default=4, help="ensemble members per shape")
    ap.add_argument("--dt", type=float, default=1e-7)
    ap.add_argument("--tune-steps", type=int, default=3, help="timed steps per autotuner pass")
    ap.add_argument("--steps", type=int, default=20, help="timed steps of each before/after measurement")
    ap.add_argument("--out", default=None, help="also write the results as JSON lines here")
    args = ap.parse_args()
    shapes = SERVED if not args.shapes else [tuple(int(v) for v in s.split("x")) for s in args.shapes.split(",")]
    fout = open(args.out, "w") if args.out else None
    entries = []
    for Nr, Nm in shapes:
        prev = fdns.get_tuning(Nr, Nm)
        runs = members(Nr, Nm, args.runs, args.dt)
        ens = fdns.Ensemble(runs)
        try:
            fdns.set_tuning(Nr, Nm)   # the plain choice
            t_plain = ms_per_step(ens, args.steps)
            chosen, _ = ens.autotune(steps=args.tune_steps)
            t_tuned = ms_per_step(ens, args.steps)
        finally:
            ens.close()
            for r in runs:
                r.close()
            fdns.set_tuning(Nr, Nm, **prev)   # the script reports; it stores nothing
        if t_tuned > t_plain:   # keep what measured faster
            chosen = fdns.make_tuning().as_dict()
        rec = dict(Nr=Nr, Nm=Nm, runs=args.runs, ms_plain=round(t_plain, 4), ms_tuned=round(t_tuned, 4),
                   gain=round(t_plain / t_tuned, 3), chosen=chosen)
        print(json.dumps(rec), flush=True)
        if fout:
            fout.write(json.dumps(rec) + "\n")
            fout.flush()
        entries.append(c_entry(Nr, Nm, chosen))
    print("\n".join(entries), flush=True)


if __name__ == "__main__":
    main()
