"""Checkpoint and snapshot files, and the restart start-up (CPU side).

* The oracle's restart from the states of K consecutive checkpoints
  (get_checkpoint_data, init.F90:447-559; steptime.F90:87-193) continues an
  uninterrupted run: the histories it rebuilds are functions of the past
  states, so a restart from the right checkpoints in the right slots
  reproduces the run to rounding, and any other placement does not.
* The oracle's history arrays, in the checkpoint's layout (order index
  fastest, slot 1 the newest), hold the states the run passed through.
* The file readers (fdns.files) parse records written by the Fortran runtime
  itself (flang's unformatted sequential I/O, the reference's writer).
* The Fortran driver refuses a restart whose checkpoint is missing or has
  another resolution, before any device work.
"""
from __future__ import annotations

import os
import subprocess

import numpy as np
import pytest

from conftest import rel

Nr, Nm, ETA = 17, 8, 0.35
FIELDS = ("tFR", "omgFR", "urFR", "upFR")


def _states(orc, scheme, nsteps, dt=1e-4, Ra=1e5):
    """An uninterrupted oracle run: the state and time after every step."""
    A = orc.Run(Nr, Nm, Ra, 1.0, dt, scheme=scheme, eta=ETA)
    A.set_state(*orc.seed_state(Nr, Nm, ETA, 7, amp=1e-2))
    states, times = [], []
    for _ in range(nsteps):
        A.step(1)
        g = A.get_state()
        states.append(tuple(g[f] for f in FIELDS))
        times.append(A.diag()["t"])
    return A, states, times


def _err(got, ref):
    return max(rel(got[f], ref[f]) for f in FIELDS)


@pytest.mark.parametrize("scheme", ["CNAB2", "SBDF2", "SBDF3", "SBDF4"])
def test_oracle_restart_continues_run(orc, scheme):
    ni = orc.scheme_info(*orc.SCHEMES[scheme])["n_imp"]
    A, states, times = _states(orc, scheme, 8)
    A.step(5)
    ref = A.get_state()
    B = orc.Run(Nr, Nm, 1e5, 1.0, 1e-4, scheme=scheme, eta=ETA, restart=True)
    B.restart(states[-ni:], times[-1])
    B.step(5)
    assert abs(B.diag()["t"] - A.diag()["t"]) < 1e-15
    assert _err(B.get_state(), ref) < 1e-12
    # the same checkpoints in the wrong slots (oldest and newest swapped)
    swapped = [states[-1]] + states[-ni + 1:-1] + [states[-ni]]
    C = orc.Run(Nr, Nm, 1e5, 1.0, 1e-4, scheme=scheme, eta=ETA, restart=True)
    C.restart(swapped, times[-1])
    C.step(5)
    assert _err(C.get_state(), ref) > 1e-6


def test_oracle_restart_cnab3_explicit_slots(orc):
    """CN/AB3 reads two checkpoints (init.F90:464-486) into explicit slots
    n_exp - j = 3, 2 (steptime.F90:88-106): slot 3 holds the n-1 state's
    term, slot 2 the n state's, slot 1 stays empty until the first step."""
    _, states, times = _states(orc, "CNAB3", 4)
    B = orc.Run(Nr, Nm, 1e5, 1.0, 1e-4, scheme="CNAB3", eta=ETA, restart=True)
    B.restart(states[-2:], times[-1])
    hb = B.history(2, 3)
    C = orc.Run(Nr, Nm, 1e5, 1.0, 1e-4, scheme="CNAB2", eta=ETA, restart=True)
    C.restart(states[-2:], times[-1])
    hc = C.history(2, 2)
    assert np.array_equal(hb["exp_temp"][2], hc["exp_temp"][1])
    assert np.array_equal(hb["exp_temp"][1], hc["exp_temp"][0])
    assert not hb["exp_temp"][0].any() and not hb["exp_ub"][0].any()
    assert np.array_equal(hb["imp_temp"], hc["imp_temp"])


@pytest.mark.parametrize("scheme", ["CNAB2", "SBDF3", "SBDF4"])
def test_oracle_history_layout(orc, scheme):
    """After S steps, rhs_imp_temp(k) holds the temperature the run had after
    step S - k (the value the step S + 1 - k started from), rhs_imp_uphi_bar(k)
    the mean flow, rhs_imp_vort(k) the vorticity of modes m >= 1; the m = 0
    vorticity rows are never written by a step (fourierloop.F90:145-171)."""
    info = orc.scheme_info(*orc.SCHEMES[scheme])
    ni, ne = info["n_imp"], info["n_exp"]
    S = 7
    A, states, _ = _states(orc, scheme, S)
    h = A.history(ni, ne)
    assert h["imp_temp"].shape == (ni, Nm + 1, Nr) and h["exp_temp"].shape == (ne, Nm + 1, Nr)
    assert h["imp_ub"].shape == (ni, Nr)
    for k in range(ni):
        st = states[S - 2 - k]
        assert np.array_equal(h["imp_temp"][k], st[0])
        assert np.array_equal(h["imp_vort"][k][1:], st[1][1:])
        assert np.array_equal(h["imp_ub"][k], st[3][0])
        assert not h["imp_vort"][k][0].any()
    for k in range(ne):
        assert np.abs(h["exp_temp"][k]).max() > 0


def test_oracle_restart_refuses_wrong_count(orc):
    _, states, times = _states(orc, "SBDF3", 3)
    B = orc.Run(Nr, Nm, 1e5, 1.0, 1e-4, scheme="SBDF3", eta=ETA, restart=True)
    with pytest.raises(ValueError):
        B.restart(states[-2:], times[-1])
    B.restart(states[-1:], times[-1])          # l_imexrk_started = .FALSE.: one checkpoint
    h = B.history(3, 3)
    assert not any(h[k].any() for k in h)


# ------------------------------------------------------------------ files
FWRITER = """program w
   implicit none
   integer, parameter :: dp = kind(1.0d0)
   integer, parameter :: Nm = {Nm}, Nr = {Nr}, ni = 3, ne = 2, nmax = 3
   complex(dp) :: f(Nm+1, Nr, 4), hi(ni, Nm+1, Nr), he(ne, Nm+1, Nr), ui(ni, Nr), ue(ne, Nr)
   real(dp) :: dta(nmax), rad(Nr)
   integer :: m, r, k, u
   do r = 1, Nr
      do m = 0, Nm
         do k = 1, 4
            f(m+1, r, k) = cmplx(m + 100*r + 10000*k, -m - 0.5_dp*r, kind=dp)
         end do
         do k = 1, ni
            hi(k, m+1, r) = cmplx(k + 10*m, r, kind=dp)
         end do
         do k = 1, ne
            he(k, m+1, r) = cmplx(-k, m + 1000*r, kind=dp)
         end do
      end do
      rad(r) = 1.0_dp + r
      do k = 1, ni
         ui(k, r) = cmplx(k*r, 0, kind=dp)
      end do
      do k = 1, ne
         ue(k, r) = cmplx(0, k*r, kind=dp)
      end do
   end do
   dta = 0.25_dp
   open (newunit=u, file="checkpoint_00007", action="write", status="replace", form="unformatted")
   write (u) 0.25_dp, 1.75_dp, Nm, Nr, ni, ne, nmax
   write (u) dta
   write (u) f(:, :, 1)
   write (u) f(:, :, 2)
   write (u) f(:, :, 3)
   write (u) f(:, :, 4)
   write (u) hi
   write (u) he
   write (u) hi
   write (u) he
   write (u) ui
   write (u) ue
   close (u)
   open (newunit=u, file="snapshot_plot_00003", action="write", status="replace", form="unformatted")
   write (u) 1.0e5_dp, 0.7_dp, 0.35_dp
   write (u) 0.25_dp, 1.75_dp
   write (u) Nm, Nr
   write (u) rad
   write (u) f(:, :, 1)
   write (u) f(:, :, 2)
   write (u) f(:, :, 3)
   write (u) f(:, :, 4)
   close (u)
end program w
"""


def _flang():
    from fdns import build
    if not os.path.exists(build.FC):
        pytest.skip("no Fortran compiler")
    return build.FC


def test_readers_parse_fortran_runtime_records(tmp_path):
    """Files written by flang's own unformatted I/O, in the record layout of
    store_checkpoint and store_snapshot, read back value for value."""
    from fdns import files
    src = tmp_path / "w.f90"
    src.write_text(FWRITER.format(Nm=Nm, Nr=Nr))
    exe = tmp_path / "w"
    subprocess.run([_flang(), str(src), "-o", str(exe)], check=True, cwd=tmp_path, capture_output=True)
    subprocess.run([str(exe)], check=True, cwd=tmp_path)
    m = np.arange(Nm + 1)[:, None]
    r = np.arange(1, Nr + 1)[None, :]
    want = [(m + 100 * r + 10000 * k) + 1j * (-m - 0.5 * r) for k in range(1, 5)]
    c = files.read_checkpoint(tmp_path / "checkpoint_00007")
    assert (c["dt"], c["time"], c["Nm"], c["Nr"], c["n_imp"], c["n_exp"], c["n_max"]) == (0.25, 1.75, Nm, Nr, 3, 2, 3)
    assert np.array_equal(c["dt_array"], [0.25] * 3)
    for name, w in zip(FIELDS, want):
        assert np.array_equal(c[name], w)
    k = np.arange(1, 4)[:, None, None]
    assert np.array_equal(c["imp_temp"], (k + 10 * m[None]) + 1j * r[None])
    assert np.array_equal(c["imp_vort"], c["imp_temp"])
    k2 = np.arange(1, 3)[:, None, None]
    assert np.array_equal(c["exp_temp"], -k2 + 1j * (m[None] + 1000 * r[None]) + 0 * k2)
    assert np.array_equal(c["imp_ub"], np.arange(1, 4)[:, None] * np.arange(1, Nr + 1)[None, :] + 0j)
    assert np.array_equal(c["exp_ub"], 1j * np.arange(1, 3)[:, None] * np.arange(1, Nr + 1)[None, :])
    s = files.read_snapshot(tmp_path / "snapshot_plot_00003")
    assert (s["Ra"], s["Pr"], s["eta"], s["dt"], s["time"], s["Nm"], s["Nr"]) == (1e5, 0.7, 0.35, 0.25, 1.75, Nm, Nr)
    assert np.array_equal(s["radius"], 1.0 + np.arange(1, Nr + 1))
    for name, w in zip(FIELDS, want):
        assert np.array_equal(s[name], w)
    assert "omgFR_check" not in s


def test_reader_rejects_damaged_file(tmp_path):
    from fdns import files
    p = tmp_path / "bad"
    p.write_bytes(np.array([8], "<i4").tobytes() + b"\0" * 8 + np.array([9], "<i4").tobytes())
    with pytest.raises(ValueError, match="markers differ"):
        files.read_records(p)
    p.write_bytes(np.array([64], "<i4").tobytes() + b"\0" * 8)
    with pytest.raises(ValueError, match="bad record length"):
        files.read_records(p)


# ------------------------------------------------------------------ driver
def _restart_input(tmp_path):
    from test_driver import write_input
    inp = write_input(tmp_path / "input.txt", Nr=Nr, Nm=Nm, steps=2, stype="IMEX", imp="BDF3", exp="AB3")
    text = inp.read_text().replace("l_restart=.FALSE.", "l_restart=.TRUE.")
    text = text.replace("l_imexrk_started = .FALSE.", "l_imexrk_started = .TRUE.")
    text = text.replace("n_restart_point=1", "n_restart_point=7")
    inp.write_text(text)
    return inp


def test_driver_restart_needs_checkpoints(tmp_path):
    from test_driver import driver
    inp = _restart_input(tmp_path)
    p = subprocess.run([driver(), str(inp)], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    # BDF3 with l_imexrk_started reads checkpoints 5, 6 and 7
    assert p.returncode == 2 and "checkpoint_00005" in p.stderr


def test_driver_restart_rejects_other_resolution(tmp_path):
    """A checkpoint of another Nm_max: the driver stops before device work."""
    from test_driver import driver
    src = tmp_path / "w.f90"
    src.write_text(FWRITER.format(Nm=Nm + 4, Nr=Nr))
    subprocess.run([_flang(), str(src), "-o", str(tmp_path / "w")], check=True, cwd=tmp_path, capture_output=True)
    for k in (5, 6, 7):
        subprocess.run([str(tmp_path / "w")], check=True, cwd=tmp_path)
        os.replace(tmp_path / "checkpoint_00007", tmp_path / f"checkpoint_{k:05d}.tmp")
    for k in (5, 6, 7):
        os.replace(tmp_path / f"checkpoint_{k:05d}.tmp", tmp_path / f"checkpoint_{k:05d}")
    inp = _restart_input(tmp_path)
    p = subprocess.run([driver(), str(inp)], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert p.returncode == 2 and "change of resolution" in p.stderr
