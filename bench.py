#!/usr/bin/env python3
"""Throughput of the MI355X FP64 time-stepping core of FAlCon-DNS.

Headline workload (default): the batched ensemble of 64 independent annulus
runs, Nr_max=97, Nm_max=256, Ra log-uniform in [1e5, 1e7], Pr in {0.7, 1, 7},
schemes CNAB2 / SBDF3 / ARS443 (run k takes scheme k mod 3), dt=1e-6, each run
seeded.  One "step" advances all 64 runs by one time step; the metric is
mode-steps/s = sum over runs of (Nm+1) per step / second.  The same 64 runs are
also advanced one at a time on the plain single-run path, and the ratio is
reported.

Prints one JSON line (see the driver contract in the repository docs).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "falcon-dns_amd"))

import fdns  # noqa: E402

METRIC = "mode-steps/s batched vs plain (FP64, ≤1e-10 rel); p50/p99 job latency"
HBM_PEAK_GBS = 8000.0


def ensemble_members(nruns=64, Nr=97, Nm=256, dt=1e-6, base_seed=2000):
    """The 64-run mixed ensemble (Ra log-uniform, Pr and scheme cycled)."""
    rng = np.random.default_rng(base_seed)
    u = rng.random(nruns)
    schemes = ("CNAB2", "SBDF3", "ARS443")
    prs = (0.7, 1.0, 7.0)
    out = []
    for k in range(nruns):
        sch = schemes[k % 3]
        out.append(dict(Nr=Nr, Nm=Nm, Ra=float(10.0 ** (5.0 + 2.0 * u[k])), Pr=prs[(k // 3) % 3], dt=dt,
                        scheme=sch, seed=base_seed + k, restart=(sch == "SBDF3")))
    return out


def make_run(m, n_KE=0):
    r = fdns.Run(m["Nr"], m["Nm"], m["Ra"], m["Pr"], m["dt"], m["scheme"], restart=m["restart"], n_KE=n_KE)
    r.seed(m["seed"])
    return r


# ---------------------------------------------------------------- roofline
# Algorithmic HBM bytes of each kernel class per batched step.  A radial vector
# is Nr complex (16*Nr bytes); a factor set streams AT (Nr^2) and AF (4 Nr^2)
# FP64 entries per mode once per launch, however many runs share it.
def rk_stage_vectors(s, S=5, fa_prev=True, need_fd=True):
    """Vectors one IMEXRK stage kernel moves per mode at stage s (1..S-1):
    reads Y0 of T and omega (2) and the earlier F_d, F_a (T and omega) and
    buoyancy terms (5s), the four products when F_a(s-1) is formed (4);
    writes T, omega, u_phi, u_r, psi (5), the buoyancy (1), F_a(s-1) (2),
    F_d(s) (2) and, at the last stage, Y0 (2)."""
    v = 2 + 5 * s + 5 + 1
    if fa_prev:
        v += 4 + 2
    if need_fd:
        v += 2
    if s == S - 1:
        v += 2
    return v


def imex_vectors(scheme):
    """Vectors one IMEX mode update moves: the four products, T and omega and
    their coefficients (4), history reads beyond order 0 (CNAB2: 2 explicit;
    SBDF3: 2 implicit + 2 explicit, twice for T and omega), history writes (4),
    new T, omega, u_phi, u_r, psi and the two coefficient vectors (7)."""
    hist = {"CNAB2": 2, "SBDF3": 8}[scheme]
    return 4 + 4 + hist + 4 + 7


def algorithmic_bytes_per_step(members):
    Nr, Nm1 = members[0]["Nr"], members[0]["Nm"] + 1
    vecb = 16 * Nr
    fset = 8 * 5 * Nr * Nr * Nm1            # AT + AF of every mode
    ars = [m for m in members if m["scheme"] == "ARS443"]
    imx = [m for m in members if m["scheme"] != "ARS443"]
    out = {"nonlin": 0.0, "imex": 0.0, "rk_stage": 0.0, "other": 0.0}
    # nonlinear pass: 4 spectral fields in and 4 out per radial level
    nl_per_run = 8 * 16 * Nm1 * Nr
    out["nonlin"] = nl_per_run * (len(imx) + 4 * len(ars))
    if imx:
        sets = {(m["scheme"], m["Pr"]) for m in imx}
        out["imex"] = len(sets) * fset + sum(imex_vectors(m["scheme"]) for m in imx) * Nm1 * vecb
    if ars:
        sets = {m["Pr"] for m in ars}
        per_stage = [len(sets) * fset + len(ars) * Nm1 * vecb * rk_stage_vectors(s, need_fd=(s < 4))
                     for s in (1, 2, 3, 4)]
        out["rk_stage"] = float(sum(per_stage))
    return out


def cpu_baseline(members, steps, threads):
    """The oracle (C restatement, OpenMP over modes like the reference) on a
    bounded sample: one run of each scheme of the ensemble."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle as O  # checker only: the CPU baseline leg
    picked = {}
    for m in members:
        picked.setdefault(m["scheme"], m)
    total_modes = 0
    t_total = 0.0
    for sch, m in picked.items():
        o = O.Run(m["Nr"], m["Nm"], m["Ra"], m["Pr"], m["dt"], sch, restart=m["restart"], nthreads=threads)
        o.set_state(*O.seed_state(m["Nr"], m["Nm"], 0.35, m["seed"]))
        o.step(1)                       # factorisation and first step, untimed
        t0 = time.perf_counter()
        o.step(steps)
        t_total += time.perf_counter() - t0
        total_modes += (m["Nm"] + 1) * steps
    return dict(value=total_modes / t_total, unit="mode-steps/s", cores=threads, kind="port",
                sample=f"oracle C restatement, one run each of {sorted(picked)} at Nr={members[0]['Nr']} "
                       f"Nm={members[0]['Nm']}, {steps} timed steps each after one untimed step")


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--runs", type=int, default=64)
    ap.add_argument("--plain-steps", type=int, default=None, help="steps per run on the plain path (default: --steps)")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=10)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); ensembles shard by run with no "
              "communication, so N GPUs would each run their own share.", file=sys.stderr)
        sys.exit(2)

    members = ensemble_members(args.runs)
    Nr, Nm = members[0]["Nr"], members[0]["Nm"]
    Nm1 = Nm + 1
    runs = [make_run(m) for m in members]
    ens = fdns.Ensemble(runs)
    ens.step(max(args.warmup, 1), graph=True)

    # ---------------- timed region: batched steps, HIP events on every launch
    t0 = time.perf_counter()
    wall_ms, kern = ens.step_timed(args.steps)
    t1 = time.perf_counter()
    elapsed = t1 - t0
    mode_steps = len(members) * Nm1 * args.steps
    value = mode_steps / elapsed

    # the same batched step replayed from a captured hipGraph (not the headline)
    tg0 = time.perf_counter()
    ens.step(args.steps, graph=True)
    graph_elapsed = time.perf_counter() - tg0

    # ---------------- dominant kernel and its roofline
    per_step = algorithmic_bytes_per_step(members)
    dom = max(kern, key=lambda c: kern[c][0])
    dom_ms, dom_n = kern[dom]
    launches_per_step = dom_n / args.steps
    bytes_launch = per_step[dom] / max(launches_per_step, 1e-30)
    avg_ms = dom_ms / max(dom_n, 1)
    achieved = bytes_launch / (avg_ms * 1e-3) / 1e9
    roofline = dict(kernel=dom, bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=round(achieved / HBM_PEAK_GBS, 4), traffic=None, bytes_per_launch=int(bytes_launch),
                    avg_launch_ms=round(avg_ms, 4), launches=dom_n,
                    per_class_GBs={c: round(per_step[c] * args.steps / (kern[c][0] * 1e-3) / 1e9, 1)
                                   for c in per_step if kern[c][0] > 0})

    # ---------------- plain path: the same runs one at a time
    plain = None
    if not args.no_plain:
        ens.close()
        psteps = args.plain_steps or args.steps
        fresh = [make_run(m) for m in members]
        t_plain = 0.0
        for r in fresh:
            r.step(max(args.warmup, 1))
            tp0 = time.perf_counter()
            r.step(psteps)
            t_plain += time.perf_counter() - tp0
        plain_value = len(members) * Nm1 * psteps / t_plain
        plain = dict(value=plain_value, unit="mode-steps/s", steps_per_run=psteps, seconds=round(t_plain, 4))
        for r in fresh:
            r.close()

    cpu = None
    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        cpu = cpu_baseline(members, args.cpu_steps, threads)

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "mode-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded initial temperature perturbations, build-defined)",
        "config": {"workload": "batched ensemble of 64 runs, Nr=97 Nm=256, Ra log-uniform 1e5..1e7, Pr {0.7,1,7}, "
                               "schemes CNAB2/SBDF3/ARS443, dt=1e-6",
                   "runs": len(members), "Nr": Nr, "Nm": Nm, "dt": members[0]["dt"],
                   "schemes": {s: sum(1 for m in members if m["scheme"] == s) for s in ("CNAB2", "SBDF3", "ARS443")}},
        "graph_value": mode_steps / graph_elapsed,
        "kernel_ms": {k: round(v[0], 3) for k, v in kern.items()},
        "kernel_launches": {k: v[1] for k, v in kern.items()},
        "plain": plain,
        "speedup_vs_plain": (value / plain["value"]) if plain else None,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
