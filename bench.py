#!/usr/bin/env python3
"""Throughput of the MI355X FP64 time-stepping core of FAlCon-DNS.

Headline workload (default): the batched ensemble of 64 independent annulus
runs, Nr_max=97, Nm_max=256, Ra log-uniform in [1e5, 1e7], Pr in {0.7, 1, 7},
schemes CNAB2 / SBDF3 / ARS443 (run k takes scheme k mod 3), dt=1e-6, each run
seeded.  One "step" advances all 64 runs by one time step; the metric is
mode-steps/s = sum over runs of (Nm+1) per step / second.  The same 64 runs are
also advanced one at a time on the plain single-run path, and the ratio is
reported.  A served stream of uneven jobs (Poisson arrivals into the
continuous-batching pool) reports p50/p99 time to completion.

Prints one JSON line (see the driver contract in the repository docs).
"""
from __future__ import annotations

import argparse
import json
import os
import re
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


def make_run(m, n_KE=0, solver="auto"):
    r = fdns.Run(m["Nr"], m["Nm"], m["Ra"], m["Pr"], m["dt"], m["scheme"], restart=m["restart"], n_KE=n_KE,
                 solver=solver)
    r.seed(m["seed"])
    return r


# ---------------------------------------------------------------- roofline
# Algorithmic HBM bytes of each kernel class per batched step: the radial
# vectors each kernel must read and write (16*Nr bytes per complex vector) and
# the operator data each launch must stream once per distinct set, however many
# runs share it.
def rk_stage_vectors(s, S=5, fa_prev=True, need_fd=True):
    """Vectors one IMEXRK stage kernel moves per mode at stage s (1..S-1):
    reads Y0 of T and omega (2), the earlier F_d, F_a (T and omega) and buoyancy
    terms (5s), the four products when F_a(s-1) is formed (4); writes T, omega,
    u_phi, u_r, psi (5), the buoyancy (1), F_a(s-1) of T and omega (2), F_d(s)
    (2) and, at the last stage, Y0 (2)."""
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
    their coefficient vectors (4), history reads beyond the newest level
    (CNAB2: 2 explicit; SBDF3: 2 implicit + 2 explicit for T and omega), history
    writes (4), new T, omega, u_phi, u_r, psi and two coefficient vectors (7)."""
    hist = {"CNAB2": 2, "SBDF3": 8}[scheme]
    return 4 + 4 + hist + 4 + 7


def op_set_bytes(N, Nm, kind, lu):
    """Bytes of one resident operator set (as the library stores them)."""
    if kind == "T":
        return 8 * N * N * (Nm + 1) + (4 * N * (Nm + 1) if lu else 4 * N)
    if lu:
        return 8 * (4 * N * N * Nm + N * N) + 4 * (2 * N * Nm + N)
    return 8 * (2 * N * (N - 2) * Nm + N * N)


def op_sets(members, lu):
    """Distinct (T, F) operator sets of a group of runs."""
    wt = {"CNAB2": 0.5, "SBDF3": 6.0 / 11.0, "ARS443": 0.5}
    irk = {"CNAB2": 0, "SBDF3": 0, "ARS443": 1}
    T = {(wt[m["scheme"]] * m["dt"], m["Pr"]) for m in members}
    F = {(wt[m["scheme"]] * m["dt"], irk[m["scheme"]] if lu else 0) for m in members}
    return len(T), len(F)


def algorithmic_bytes_per_step(members, lu):
    Nr, Nm = members[0]["Nr"], members[0]["Nm"]
    Nm1 = Nm + 1
    vecb = 16 * Nr
    ars = [m for m in members if m["scheme"] == "ARS443"]
    imx = [m for m in members if m["scheme"] != "ARS443"]
    out = {"nonlin": 0.0, "imex": 0.0, "rk_stage": 0.0, "other": 0.0}
    # nonlinear pass: 4 spectral fields in and 4 out per radial level
    nl_per_run = 8 * 16 * Nm1 * Nr
    out["nonlin"] = float(nl_per_run * (len(imx) + 4 * len(ars)))
    bT, bF = op_set_bytes(Nr, Nm, "T", lu), op_set_bytes(Nr, Nm, "F", lu)
    if imx:
        nT, nF = op_sets(imx, lu)
        out["imex"] = float(nT * bT + nF * bF + sum(imex_vectors(m["scheme"]) for m in imx) * Nm1 * vecb)
    if ars:
        nT, nF = op_sets(ars, lu)
        out["rk_stage"] = float(sum(nT * bT + nF * bF + len(ars) * Nm1 * vecb * rk_stage_vectors(s, need_fd=(s < 4))
                                    for s in (1, 2, 3, 4)))
    return out


def measured_traffic(kernel_class):
    """HBM bytes per launch of the kernel class from the committed rocprofv3
    PMC summary (FETCH_SIZE + WRITE_SIZE, scaled by the copy calibration)."""
    path = os.path.join(REPO, "profiles", "pmc_traffic_c2_ensemble.md")
    if not os.path.exists(path):
        return None
    m = re.search(rf"^traffic_bytes_per_launch\s+{kernel_class}\s+(\d+)", open(path).read(), flags=re.M)
    return int(m.group(1)) if m else None


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


# ---------------------------------------------------------------- serving
def serve_jobs(njobs, seed=3000, step_scale=1.0):
    """The served stream: Nr in {33, 65, 129, 257}, Nm in {64, ..., 768}, Ra
    log-uniform in [1e5, 1e8], Pr in {0.7, 1, 7}, scheme in {CNAB2, SBDF3,
    ARS443, BPR353}, 500-5000 steps (times step_scale), dt = 1e-7."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(njobs):
        out.append(dict(Nr=int(rng.choice([33, 65, 129, 257])), Nm=int(rng.choice([64, 128, 256, 512, 768])),
                        Ra=float(10.0 ** rng.uniform(5.0, 8.0)), Pr=float(rng.choice([0.7, 1.0, 7.0])),
                        scheme=str(rng.choice(["CNAB2", "SBDF3", "ARS443", "BPR353"])), dt=1e-7,
                        steps=max(1, int(round(int(rng.integers(500, 5001)) * step_scale))), seed=seed + k))
    return out


def serve(njobs, rate, budget, step_scale, seed=3000):
    """Poisson arrivals (rate jobs/s; rate <= 0: all at t = 0) into the pool; every pool step admits
    what fits, advances all resident jobs together and retires finished ones.
    Returns the time-to-completion percentiles and the throughput."""
    jobs = serve_jobs(njobs, seed, step_scale)
    rng = np.random.default_rng(seed + 1)
    if rate > 0:
        arrivals = np.cumsum(rng.exponential(1.0 / rate, size=njobs))
    else:                               # every job at t = 0: the pool's saturation throughput
        arrivals = np.zeros(njobs)
    P = fdns.Pool(budget)
    ids, t_arr, t_done = {}, {}, {}
    nxt = 0
    max_res = 0
    t0 = time.perf_counter()
    while len(t_done) < njobs:
        now = time.perf_counter() - t0
        while nxt < njobs and arrivals[nxt] <= now:
            j = jobs[nxt]
            ids[nxt] = P.submit(j["Nr"], j["Nm"], j["Ra"], j["Pr"], j["dt"], j["scheme"], j["steps"], j["seed"],
                                restart=j["scheme"] == "SBDF3")
            t_arr[nxt] = arrivals[nxt]
            nxt += 1
        pending = [k for k in ids if k not in t_done]
        if not pending:
            time.sleep(max(0.0, arrivals[nxt] - now))
            continue
        active, waiting = P.step()
        max_res = max(max_res, active)
        now = time.perf_counter() - t0
        for k in pending:
            if P.poll(ids[k])[0] == 2:
                t_done[k] = now
                P.release(ids[k])
    wall = time.perf_counter() - t0
    P.close()
    lat = np.array([t_done[k] - t_arr[k] for k in range(njobs)])
    mode_steps = sum((j["Nm"] + 1) * j["steps"] for j in jobs)
    return dict(jobs=njobs, rate_jobs_per_s=rate, p50_s=float(np.percentile(lat, 50)),
                p99_s=float(np.percentile(lat, 99)), mean_s=float(lat.mean()), wall_s=wall,
                throughput=mode_steps / wall, unit="mode-steps/s", max_resident=max_res, budget_bytes=int(budget),
                dt=1e-7, step_scale=step_scale,
                workload="Poisson arrivals; Nr {33,65,129,257}, Nm {64..768}, Ra 1e5..1e8, Pr {0.7,1,7}, "
                         "CNAB2/SBDF3/ARS443/BPR353, 500-5000 steps x step_scale, dt 1e-7")


def calibrate_copy(nbytes=1 << 30, reps=5):
    """A plain device copy of known size, for scaling the PMC byte counters."""
    import torch
    a = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": nbytes, "copy_bytes_written": nbytes, "reps": reps}))


def main():
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--runs", type=int, default=64)
    ap.add_argument("--solver", default="auto", choices=["auto", "lu", "inverse"])
    ap.add_argument("--plain-steps", type=int, default=None, help="steps per run on the plain path (default: --steps)")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=10)
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--serve-jobs", type=int, default=32)
    # default rate: ~75 % of the pool's measured saturation rate for this stream
    # (32 jobs at step scale 0.1 all submitted at t = 0 finish in ~5.4 s on one
    # MI355X, i.e. ~5.9 jobs/s; --serve-rate 0 repeats that measurement)
    ap.add_argument("--serve-rate", type=float, default=4.5, help="Poisson arrival rate, jobs/s (0: all at t = 0)")
    ap.add_argument("--serve-step-scale", type=float, default=0.1)
    ap.add_argument("--serve-budget-gib", type=float, default=64.0)
    ap.add_argument("--calibrate-copy", action="store_true", help="only run a 1 GiB device copy (PMC calibration)")
    ap.add_argument("--no-graph", action="store_true", help="no hipGraph capture or replay (eager launches only)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); ensembles shard by run with no "
              "communication, so N GPUs would each run their own share.", file=sys.stderr)
        sys.exit(2)
    if args.calibrate_copy:
        calibrate_copy()
        return

    members = ensemble_members(args.runs)
    Nr, Nm = members[0]["Nr"], members[0]["Nm"]
    Nm1 = Nm + 1
    lu = args.solver == "lu" or (args.solver == "auto" and Nr > 129)
    runs = [make_run(m, solver=args.solver) for m in members]
    ens = fdns.Ensemble(runs)
    ens.step(max(args.warmup, 1), graph=not args.no_graph)

    # ---------------- timed region: batched steps, HIP events on every launch
    t0 = time.perf_counter()
    wall_ms, kern = ens.step_timed(args.steps)
    t1 = time.perf_counter()
    elapsed = t1 - t0
    mode_steps = len(members) * Nm1 * args.steps
    value = mode_steps / elapsed

    # the same batched step replayed from a captured hipGraph
    graph_value = None
    if not args.no_graph:
        tg0 = time.perf_counter()
        ens.step(args.steps, graph=True)
        graph_value = mode_steps / (time.perf_counter() - tg0)

    # ---------------- dominant kernel and its roofline
    per_step = algorithmic_bytes_per_step(members, lu)
    dom = max(kern, key=lambda c: kern[c][0])
    dom_ms, dom_n = kern[dom]
    launches_per_step = dom_n / args.steps
    bytes_launch = per_step[dom] / max(launches_per_step, 1e-30)
    avg_ms = dom_ms / max(dom_n, 1)
    achieved = bytes_launch / (avg_ms * 1e-3) / 1e9
    roofline = dict(kernel=dom, bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=round(achieved / HBM_PEAK_GBS, 4), traffic=measured_traffic(dom),
                    bytes_per_launch=int(bytes_launch), avg_launch_ms=round(avg_ms, 4), launches=dom_n,
                    per_class_GBs={c: round(per_step[c] * args.steps / (kern[c][0] * 1e-3) / 1e9, 1)
                                   for c in per_step if kern[c][0] > 0},
                    # the kernel classes of a step overlap on parallel stream lanes, so
                    # each launch shares the chip; whole-step algorithmic bytes / wall time:
                    step_GBs=round(sum(per_step.values()) / (elapsed / args.steps) / 1e9, 1),
                    step_frac=round(sum(per_step.values()) / (elapsed / args.steps) / 1e9 / HBM_PEAK_GBS, 4))

    # ---------------- plain path: the same runs one at a time
    plain = None
    ens.close()
    for r in runs:
        r.close()
    if not args.no_plain:
        psteps = args.plain_steps or args.steps
        fresh = [make_run(m, solver=args.solver) for m in members]
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

    served = None
    if not args.no_serve:
        served = serve(args.serve_jobs, args.serve_rate, int(args.serve_budget_gib * (1 << 30)),
                       args.serve_step_scale)

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
                   "runs": len(members), "Nr": Nr, "Nm": Nm, "dt": members[0]["dt"], "solver": args.solver,
                   "solve_mode": "lu" if lu else "inverse",
                   "schemes": {s: sum(1 for m in members if m["scheme"] == s) for s in ("CNAB2", "SBDF3", "ARS443")}},
        "graph_value": graph_value,
        "kernel_ms": {k: round(v[0], 3) for k, v in kern.items()},
        "kernel_launches": {k: v[1] for k, v in kern.items()},
        "plain": plain,
        "speedup_vs_plain": (value / plain["value"]) if plain else None,
        "serving": served,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
