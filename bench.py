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
# Algorithmic HBM bytes and FP64 flops of each kernel class per batched step:
# the radial vectors (16*Nr bytes per complex vector) each kernel must read
# and write per (run, mode), the operator inverses each solve launch streams
# once per chunk of runs sharing them, and the useful MFMA work (4*N^2 flops
# per N x N real matrix applied to one complex vector).
HBM_PEAK = 8000.0          # GB/s, MI355X
MFMA_F64_PEAK = 78.6       # TFLOP/s, MI355X dense FP64 matrix
KERNEL_NAMES = {"nonlin": "k_nonlin", "imex_pre": "k_imex_pre", "imex_mid": "k_imex_mid", "imex_post": "k_imex_post",
                "rk_pre": "k_rk_pre", "rk_mid": "k_rk_mid", "rk_post": "k_rk_post", "solve_T": "k_solve<0>",
                "solve_F": "k_solve<1>", "other": "k_advance, k_imex_prep, k_diag"}
# IMEX schemes: Crank-Nicolson implicit part?, implicit and explicit history depths
IMEX_HIST = {"CNAB2": (True, 1, 2), "SBDF3": (False, 3, 3)}
# ARS443 (timeschemes.F90): S = 5 stages with an explicit first stage; the
# explicit term of stages 0..3 is read by later stages (nl_col) and so is the
# diffusion term of stages 1..3 (need_fd); stiffly accurate.
RK_INFO = {"ARS443": dict(S=5, nl_col=(1, 1, 1, 1, 0), need_fd=(0, 1, 1, 1, 0))}
SOLVE_RHS = 32             # runs per batched solve chunk


def op_chunks(members):
    """Chunks of the batched solves of one group: per distinct (T, F)
    operator set of the group's runs, ceil(runs / SOLVE_RHS)."""
    wt = {"CNAB2": 0.5, "SBDF3": 6.0 / 11.0, "ARS443": 0.5}
    T, F = {}, {}
    for m in members:
        kt = (wt[m["scheme"]] * m["dt"], m["Pr"])
        kf = wt[m["scheme"]] * m["dt"]
        T[kt] = T.get(kt, 0) + 1
        F[kf] = F.get(kf, 0) + 1
    ch = lambda d: sum((n + SOLVE_RHS - 1) // SOLVE_RHS for n in d.values())
    return ch(T), ch(F)


def kernel_model(members):
    """(bytes, flops) per batched step of each kernel class."""
    N, Nm = members[0]["Nr"], members[0]["Nm"]
    Nm1 = Nm + 1
    vec = 16 * N                          # one complex radial vector
    prod = 4 * N * N                      # one N x N product on one complex column
    B = {c: 0.0 for c in KERNEL_NAMES}
    F = {c: 0.0 for c in KERNEL_NAMES}
    imx = [m for m in members if m["scheme"] in IMEX_HIST]
    ars = [m for m in members if m["scheme"] in RK_INFO]
    opT = 8 * N * N * Nm                          # AT^{-1}, modes 1..Nm, per chunk
    opF = 8 * (2 * N * (N - 2) * Nm + N * N)      # AF^{-1} interior columns and A_uphi^{-1}, per chunk

    def solves(group, launches):
        cT, cF = op_chunks(group)
        cols = len(group) * Nm1
        B["solve_T"] += launches * (cols * 2 * vec + cT * opT)
        B["solve_F"] += launches * (cols * 3 * vec + cF * opF)
        F["solve_T"] += launches * cols * prod
        F["solve_F"] += launches * cols * 8 * N * (N - 2)

    for m in imx:
        cn, ni, ne = IMEX_HIST[m["scheme"]]
        B["nonlin"] += 8 * 16 * Nm1 * N
        B["imex_pre"] += Nm1 * vec * (4 + (0 if cn else ni - 1) + (ne - 1) + 3)
        B["imex_mid"] += Nm1 * vec * (5 + (1 if cn else ni - 1) + (ne - 1) + 5)
        B["imex_post"] += Nm1 * vec * 7
        F["imex_pre"] += Nm1 * 3 * prod
        F["imex_mid"] += Nm1 * 4 * prod
        F["imex_post"] += Nm1 * 4 * prod
    if imx:
        solves(imx, 1)
    for m in ars:
        info = RK_INFO[m["scheme"]]
        S = info["S"]
        B["nonlin"] += sum(info["nl_col"][:S - 1]) * 8 * 16 * Nm1 * N
        for s in range(1, S):
            fa = info["nl_col"][s - 1]
            last = s == S - 1
            fd = info["need_fd"][0 if last else s]
            B["rk_pre"] += Nm1 * vec * ((4 if fa else 0) + s + (s - 1 if fa else s) + 1 + (2 if fa else 0) + 1)
            B["rk_mid"] += Nm1 * vec * (1 + 3 * s + 1 + 1 + fd + 1 + 1 + last)
            B["rk_post"] += Nm1 * vec * (2 + 4 + fd + last)
            F["rk_pre"] += Nm1 * (2 * prod if fa else 0)
            F["rk_mid"] += Nm1 * (3 if fd else 1) * prod
            F["rk_post"] += Nm1 * (5 if fd else 4) * prod
    if ars:
        solves(ars, RK_INFO[ars[0]["scheme"]]["S"] - 1)
    return B, F


def kernel_table(kern, steps, B, F):
    """Per kernel class: launches, average launch time (HIP events), and the
    algorithmic GB/s and TFLOP/s that implies."""
    out = {}
    for c, (ms, n) in kern.items():
        if n == 0 or c == "other":
            continue
        per_launch_b = B[c] * steps / n
        per_launch_f = F[c] * steps / n
        avg_s = ms * 1e-3 / n
        out[c] = dict(kernel=KERNEL_NAMES[c], launches=n, avg_us=round(avg_s * 1e6, 1), total_ms=round(ms, 3),
                      bytes_per_launch=int(per_launch_b), flops_per_launch=int(per_launch_f),
                      GBs=round(per_launch_b / avg_s / 1e9, 1), TFLOPs=round(per_launch_f / avg_s / 1e12, 2),
                      hbm_frac=round(per_launch_b / avg_s / 1e9 / HBM_PEAK, 4),
                      mfma_frac=round(per_launch_f / avg_s / 1e12 / MFMA_F64_PEAK, 4))
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
SERVED_SHAPES = [(nr, nm) for nr in (33, 65, 129, 257) for nm in (64, 128, 256, 512, 768)]


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


def serve(njobs, rate, budget, step_scale, seed=3000, max_steps=16):
    """Poisson arrivals (rate jobs/s; rate <= 0: all at t = 0) into the pool; every pool call admits
    what fits, advances all resident jobs together by up to max_steps batched steps (ending at the
    first step that finishes a job) and retires finished ones.
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
        _, active, waiting = P.run(max_steps)
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
                dt=1e-7, step_scale=step_scale, max_steps_per_call=max_steps,
                workload="Poisson arrivals; Nr {33,65,129,257}, Nm {64..768}, Ra 1e5..1e8, Pr {0.7,1,7}, "
                         "CNAB2/SBDF3/ARS443/BPR353, 500-5000 steps x step_scale, dt 1e-7")


def serve_reference(njobs, steps, seed=3000):
    """The served stream's jobs advanced together as one ensemble (every job
    resident, graph replay): the batched throughput of the same shapes that
    the pool's saturation throughput is compared against."""
    jobs = serve_jobs(njobs, seed, 1.0)
    runs = []
    for j in jobs:
        r = fdns.Run(j["Nr"], j["Nm"], j["Ra"], j["Pr"], j["dt"], j["scheme"], restart=j["scheme"] == "SBDF3")
        r.seed(j["seed"])
        runs.append(r)
    ens = fdns.Ensemble(runs)
    ens.step(2, graph=True)
    t0 = time.perf_counter()
    ens.step(steps, graph=True)
    wall = time.perf_counter() - t0
    ens.close()
    for r in runs:
        r.close()
    return sum(j["Nm"] + 1 for j in jobs) * steps / wall


SINGLE = {
    # the plain single large run and the high-resolution stress run
    "c1": dict(Nr=257, Nm=768, Ra=1e8, Pr=1.0, dt=1e-7, scheme="ARS443", seed=1000, steps=20),
    "c4": dict(Nr=513, Nm=1536, Ra=1e9, Pr=1.0, dt=2e-8, scheme="BPR353", seed=4000, steps=5),
}


def single_runs():
    """The single large runs: time per step of graph-replayed steps (solver
    auto = LU order above Nr = 129) and the per-kernel-class times of the same
    steps launched with a HIP event pair around every kernel."""
    out = {}
    for name, m in SINGLE.items():
        r = fdns.Run(m["Nr"], m["Nm"], m["Ra"], m["Pr"], m["dt"], m["scheme"])
        r.seed(m["seed"])
        r.step(2)
        t0 = time.perf_counter()
        r.step(m["steps"])
        wall = time.perf_counter() - t0
        e = fdns.Ensemble([r])
        _, kern = e.step_timed(max(2, m["steps"] // 4))
        e.close()
        r.close()
        k = max(2, m["steps"] // 4)
        out[name] = dict(Nr=m["Nr"], Nm=m["Nm"], scheme=m["scheme"], steps=m["steps"],
                         ms_per_step=1e3 * wall / m["steps"], mode_steps_per_s=(m["Nm"] + 1) * m["steps"] / wall,
                         kernel_ms_per_step={c: round(v[0] / k, 4) for c, v in kern.items() if v[1] > 0})
    return out


def calibrate_copy(nbytes=1 << 30, reps=5):
    """The library's plain device copy (k_copy) of known size, for scaling
    the profiler's HBM byte counters; prints its bytes and bandwidth."""
    ms = fdns.device_copy(nbytes, reps)
    print(json.dumps({"kernel": "k_copy", "copy_bytes_read": nbytes, "copy_bytes_written": nbytes,
                      "launches": reps + 1, "avg_ms": ms, "GBs": 2 * nbytes / (ms * 1e-3) / 1e9}))


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
    ap.add_argument("--no-single", action="store_true", help="skip the single large runs (c1, c4)")
    ap.add_argument("--serve-jobs", type=int, default=32)
    # the arrival rate: by default --serve-load times the pool's saturation
    # rate for this stream, measured first (every job at t = 0, steps x 0.1)
    ap.add_argument("--serve-rate", type=float, default=None, help="Poisson arrival rate, jobs/s (0: all at t = 0)")
    ap.add_argument("--serve-load", type=float, default=0.8,
                    help="offered load as a fraction of the measured saturation throughput (when no --serve-rate)")
    ap.add_argument("--serve-step-scale", type=float, default=1.0)
    ap.add_argument("--serve-budget-gib", type=float, default=64.0)
    ap.add_argument("--serve-tuning", default="table", choices=["table", "plain"],
                    help="launch choices of the served shapes: the library's stored table, or the plain defaults")
    ap.add_argument("--calibrate-copy", action="store_true", help="only run the library's 1 GiB device copy (PMC "
                                                                     "calibration)")
    ap.add_argument("--no-graph", action="store_true", help="no hipGraph capture or replay (eager launches only)")
    ap.add_argument("--tuning", default="auto",
                    help="launch choices: 'auto' (fdns_ens_autotune on the ensemble before the warmup), 'table' "
                         "(the library's stored choice for the shape), or "
                         "nl_tpb,tile_waves,tile_stage,solve_chunk[,tile_pf[,lane_split]], or a JSON object of "
                         "fdns.set_tuning's arguments")
    ap.add_argument("--tune-steps", type=int, default=3, help="timed steps per autotuner pass")
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
    if lu:
        print("bench.py models the inverse-mode solves of the 64-run ensemble (Nr=97); --solver lu times "
              "the dgetrs path with the same byte model", file=sys.stderr)
    runs = [make_run(m, solver=args.solver) for m in members]
    ens = fdns.Ensemble(runs)
    tuning = {"mode": args.tuning}
    if args.tuning == "auto":
        t_tune = time.perf_counter()
        chosen, report = ens.autotune(args.tune_steps)
        tuning.update(chosen=chosen, report_ms_per_step=report, seconds=round(time.perf_counter() - t_tune, 3))
    elif args.tuning.startswith("{"):
        fdns.set_tuning(Nr, Nm, **json.loads(args.tuning))
    elif args.tuning != "table":
        v = [int(x) for x in args.tuning.split(",")]
        fdns.set_tuning(Nr, Nm, *v)
    tuning["used"] = fdns.get_tuning(Nr, Nm)
    ens.step(max(args.warmup, 1), graph=not args.no_graph)

    # ---------------- timed region: K batched steps, each one replay of the
    # captured hipGraph of the step (eager launches with --no-graph); the
    # library synchronises the stream at the end of every call
    t0 = time.perf_counter()
    ens.step(args.steps, graph=not args.no_graph)
    elapsed = time.perf_counter() - t0
    mode_steps = len(members) * Nm1 * args.steps
    value = mode_steps / elapsed

    # ---------------- the same K steps launched eagerly, the groups one after
    # another, with a HIP event pair around every kernel launch: per-kernel
    # durations of each kernel alone on the chip
    t1 = time.perf_counter()
    wall_ms, kern = ens.step_timed(args.steps)
    eager_value = mode_steps / (time.perf_counter() - t1)

    # ---------------- per-kernel rates; the dominant kernel and its roofline
    B, F = kernel_model(members)
    table = kernel_table(kern, args.steps, B, F)
    dom = max(table, key=lambda c: table[c]["total_ms"])
    d = table[dom]
    ai = d["flops_per_launch"] / max(d["bytes_per_launch"], 1)
    if ai > MFMA_F64_PEAK * 1e12 / (HBM_PEAK * 1e9):
        bound, achieved, peak, unit = "mfma", d["TFLOPs"], MFMA_F64_PEAK, "TFLOP/s"
    else:
        bound, achieved, peak, unit = "hbm", d["GBs"], HBM_PEAK, "GB/s"
    step_bytes = sum(B.values())
    roofline = dict(kernel=d["kernel"], bound=bound, achieved=achieved, peak=peak, unit=unit,
                    frac=round(achieved / peak, 4), traffic=measured_traffic(dom),
                    bytes_per_launch=d["bytes_per_launch"], flops_per_launch=d["flops_per_launch"],
                    avg_launch_us=d["avg_us"], launches=d["launches"],
                    # whole-step algorithmic bytes / wall time per step of the headline:
                    step_GBs=round(step_bytes / (elapsed / args.steps) / 1e9, 1),
                    step_frac=round(step_bytes / (elapsed / args.steps) / 1e9 / HBM_PEAK, 4))

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

    served = saturation = None
    if not args.no_serve:
        budget = int(args.serve_budget_gib * (1 << 30))
        if args.serve_tuning == "plain":
            for nr, nm in SERVED_SHAPES:
                fdns.set_tuning(nr, nm)
        # saturation: every job at t = 0 (shortened), against the same jobs as one ensemble
        sat = serve(args.serve_jobs, 0.0, budget, 0.1)
        ens_ref = serve_reference(args.serve_jobs, 20)
        saturation = dict(pool_throughput=sat["throughput"], ensemble_throughput=ens_ref,
                          ratio=sat["throughput"] / ens_ref, unit="mode-steps/s", step_scale=0.1,
                          max_resident=sat["max_resident"])
        rate = args.serve_rate
        if rate is None:
            jobs = serve_jobs(args.serve_jobs, step_scale=args.serve_step_scale)
            mean_job = sum((j["Nm"] + 1) * j["steps"] for j in jobs) / len(jobs)
            rate = args.serve_load * sat["throughput"] / mean_job
        served = serve(args.serve_jobs, rate, budget, args.serve_step_scale)
        served["offered_load"] = args.serve_load if args.serve_rate is None else None
        served["tuning"] = saturation["tuning"] = args.serve_tuning

    single = None if args.no_single else single_runs()

    cpu = None
    if not args.no_cpu:
        # the host cores this GPU's job is given (OMP_NUM_THREADS: 16 per GPU
        # on the MI355X hosts), else every core the process sees
        threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or (os.cpu_count() or 1)
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
        "graph": not args.no_graph,
        "eager_value": eager_value,
        "kernels": table,
        "plain": plain,
        "speedup_vs_plain": (value / plain["value"]) if plain else None,
        "serving": served,
        "serving_saturation": saturation,
        "single_runs": single,
        "tuning": tuning,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
