"""Readers of the reference's checkpoint and snapshot files.

Both are Fortran unformatted sequential files as gfortran and flang write
them: every record is its bytes framed by their length as a little-endian
int32 before and after.  Layouts:

  checkpoint_NNNNN    store_checkpoint / store_checkpoint_imexrk
                      (output.F90:359-450), read back by read_checkpoint
                      (init.F90:865-935)
  snapshot_plot_NNNNN store_snapshot / store_snapshot_imexrk
                      (output.F90:451-516), read by
                      python/annulus/snapshot.py:24-65

Pure numpy; no GPU needed.
"""
from __future__ import annotations

import numpy as np


def read_records(path):
    """The records of a Fortran unformatted sequential file, as bytes."""
    out = []
    with open(path, "rb") as f:
        data = f.read()
    pos = 0
    while pos < len(data):
        if pos + 4 > len(data):
            raise ValueError(f"{path}: truncated record marker at byte {pos}")
        n = int(np.frombuffer(data, "<i4", 1, pos)[0])
        if n < 0 or pos + 8 + n > len(data):
            raise ValueError(f"{path}: bad record length {n} at byte {pos}")
        tail = int(np.frombuffer(data, "<i4", 1, pos + 4 + n)[0])
        if tail != n:
            raise ValueError(f"{path}: record markers differ ({n} != {tail}) at byte {pos}")
        out.append(data[pos + 4:pos + 4 + n])
        pos += 8 + n
    return out


def _spec(rec, shape):
    a = np.frombuffer(rec, "<c16")
    if a.size != int(np.prod(shape)):
        raise ValueError(f"record of {a.size} complex values, expected {shape}")
    return a.reshape(shape, order="F").copy(order="F")


def read_checkpoint(path):
    """Header, dt_array, the four (Nm+1, Nr) fields and the six history
    arrays of a checkpoint file."""
    recs = read_records(path)
    if len(recs) != 12:
        raise ValueError(f"{path}: {len(recs)} records, a checkpoint has 12")
    h = recs[0]
    if len(h) != 36:
        raise ValueError(f"{path}: header of {len(h)} bytes, expected 36")
    dt, t = np.frombuffer(h, "<f8", 2, 0)
    Nm, Nr, ni, ne, nmax = (int(v) for v in np.frombuffer(h, "<i4", 5, 16))
    out = {"dt": float(dt), "time": float(t), "Nm": Nm, "Nr": Nr, "n_imp": ni, "n_exp": ne, "n_max": nmax,
           "dt_array": np.frombuffer(recs[1], "<f8").copy()}
    s = (Nm + 1, Nr)
    for k, name in enumerate(("tFR", "omgFR", "urFR", "upFR")):
        out[name] = _spec(recs[2 + k], s)
    # history arrays: n slots, order index fastest; n from the record size
    S = (Nm + 1) * Nr
    for k, name in enumerate(("imp_temp", "exp_temp", "imp_vort", "exp_vort")):
        n = len(recs[6 + k]) // (16 * S)
        out[name] = _spec(recs[6 + k], (n, Nm + 1, Nr))
    for k, name in enumerate(("imp_ub", "exp_ub")):
        n = len(recs[10 + k]) // (16 * Nr)
        out[name] = _spec(recs[10 + k], (n, Nr))
    return out


def read_snapshot(path):
    """Parameters, time, grid and the four (Nm+1, Nr) fields of a snapshot
    file (plus omgFR_check when the IMEXRK writer made it)."""
    recs = read_records(path)
    if len(recs) not in (8, 9):
        raise ValueError(f"{path}: {len(recs)} records, a snapshot has 8 or 9")
    Ra, Pr, eta = np.frombuffer(recs[0], "<f8")
    dt, t = np.frombuffer(recs[1], "<f8")
    Nm, Nr = (int(v) for v in np.frombuffer(recs[2], "<i4"))
    out = {"Ra": float(Ra), "Pr": float(Pr), "eta": float(eta), "dt": float(dt), "time": float(t),
           "Nm": Nm, "Nr": Nr, "radius": np.frombuffer(recs[3], "<f8").copy()}
    names = ["tFR", "omgFR", "urFR", "upFR"] + (["omgFR_check"] if len(recs) == 9 else [])
    for k, name in enumerate(names):
        out[name] = _spec(recs[4 + k], (Nm + 1, Nr))
    return out
