#!/usr/bin/env python3
"""Write the golden fixtures of the time-stepping core from the oracle.

The reference ships no tests, fixtures or golden vectors and cannot be built
here (it needs FFTW3 and LAPACK), so the fixtures are the oracle's output on
the seeded inputs; the oracle itself is pinned by the closed-form and
independent-library tests under tests/.  Each fixture holds the final spectral
temperature and streamfunction (Fortran (Nm+1, Nr) layout) and the
(t, KE, Nu_t, Nu_b) series.

    python tests/golden/make_golden.py            # rewrites tests/golden/*.npz
    python tests/golden/make_golden.py NAME ...   # only those fixtures
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(REPO, "oracle"))

import oracle as O  # noqa: E402

ETA = 0.35

# name -> (Nr, Nm, Ra, Pr, dt, scheme, restart, mBC, buo, steps, every, seed)
CASES = {
    # the reference's own CPU case: CNAB2, 65 x 128, Ra 1e6, 1000 steps
    "c0_cnab2_nr65_nm128": (65, 128, 1e6, 1.0, 1e-6, "CNAB2", False, "NS", "IMP", 1000, 10, 0),
    # every scheme of the registry on a small annulus
    "s_cnab2": (33, 16, 1e5, 1.0, 1e-5, "CNAB2", False, "NS", "IMP", 40, 4, 11),
    "s_cnab3": (33, 16, 1e5, 1.0, 1e-5, "CNAB3", False, "NS", "IMP", 40, 4, 12),
    "s_sbdf2": (33, 16, 1e5, 0.7, 1e-5, "SBDF2", True, "NS", "IMP", 40, 4, 13),
    "s_sbdf3": (33, 16, 1e5, 7.0, 1e-5, "SBDF3", True, "NS", "IMP", 40, 4, 14),
    "s_sbdf3_fresh": (33, 16, 1e5, 1.0, 1e-5, "SBDF3", False, "NS", "IMP", 40, 4, 15),
    "s_sbdf4": (33, 16, 1e5, 1.0, 1e-5, "SBDF4", True, "NS", "IMP", 40, 4, 16),
    "s_ars443": (33, 16, 1e5, 1.0, 1e-5, "ARS443", False, "NS", "IMP", 20, 2, 17),
    "s_ars443_exp": (33, 16, 1e5, 1.0, 1e-5, "ARS443", False, "NS", "EXP", 20, 2, 18),
    "s_bpr353": (33, 16, 1e5, 1.0, 1e-5, "BPR353", False, "NS", "IMP", 20, 2, 19),
    "s_cnab2_sf": (33, 16, 1e5, 1.0, 1e-5, "CNAB2", False, "SF", "IMP", 40, 4, 20),
    "s_ars443_sf": (33, 16, 1e5, 1.0, 1e-5, "ARS443", False, "SF", "IMP", 20, 2, 21),
    # the other IMEXRK families and the explicit RK path
    "s_ars222": (33, 16, 1e5, 1.0, 1e-5, "ARS222", False, "NS", "IMP", 20, 2, 22),
    "s_bhr553": (33, 16, 1e5, 1.0, 1e-5, "BHR553", False, "NS", "IMP", 20, 2, 23),
    "s_kc343_sf": (33, 16, 1e5, 1.0, 1e-5, "KC343", False, "SF", "IMP", 20, 2, 24),
    "s_smr333": (33, 16, 1e5, 1.0, 1e-5, "SMR333", False, "NS", "IMP", 20, 2, 25),
    "s_cb3e": (33, 16, 1e5, 1.0, 1e-5, "CB3e", False, "NS", "IMP", 20, 2, 26),
    "s_rk4": (33, 16, 1e5, 1.0, 2e-6, "RK4", False, "NS", "IMP", 20, 2, 27),
}


def run_case(Nr, Nm, Ra, Pr, dt, scheme, restart, mBC, buo, steps, every, seed, nthreads=0):
    o = O.Run(Nr, Nm, Ra, Pr, dt, scheme, mBC=mBC, buo=buo, restart=restart, nthreads=nthreads)
    o.set_state(*O.seed_state(Nr, Nm, ETA, seed))
    series = []
    done = 0
    while done < steps:
        o.step(every)
        done += every
        d = o.diag()
        series.append((d["t"], d["KE"], d["Nu_t"], d["Nu_b"]))
    s = o.get_state()
    return dict(tFR=s["tFR"], psi=s["psi"], series=np.array(series))


def main(names=None):
    for name, case in CASES.items():
        if names and name not in names:
            continue
        out = run_case(*case)
        path = os.path.join(HERE, name + ".npz")
        np.savez_compressed(path, params=np.array(repr(case)), **out)
        print(f"{name}: {os.path.getsize(path)} bytes")


if __name__ == "__main__":
    main(sys.argv[1:])
