"""Python binding of libfdns.so, the MI355X time-stepping core of FAlCon-DNS.

The C-ABI (include/fdns.h) is the product boundary; this module is a thin
ctypes mirror of it for tests, benchmarks and Python drivers.  It never falls
back to any CPU implementation: if libfdns.so is missing or no GPU is
visible, calls raise.

Spectral fields are numpy complex128 arrays of shape (Nm+1, Nr) in Fortran
order, the reference's (Nm_max+1, Nr_max) module arrays (init.F90:21-24).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(_PKG, "lib", "libfdns.so")
_lib = None


class FdnsError(RuntimeError):
    pass


class Params(C.Structure):
    """fdns_params: the namelist values that reach the time stepper."""
    _fields_ = [
        ("Nr", C.c_int32), ("Nm", C.c_int32), ("eta", C.c_double), ("Ra", C.c_double),
        ("Pr", C.c_double), ("dt", C.c_double), ("mBC", C.c_char * 8), ("buo_tscheme", C.c_char * 8),
        ("time_scheme_type", C.c_char * 16), ("time_scheme_imp", C.c_char * 16),
        ("time_scheme_exp", C.c_char * 16), ("l_restart", C.c_int32), ("n_KE", C.c_int32),
        ("solver", C.c_int32), ("l_vartimestep", C.c_int32), ("l_optimizedt", C.c_int32),
        ("CFL", C.c_double), ("dt_coef", C.c_double), ("dt_max", C.c_double), ("lagpts", C.c_int32),
    ]


# dt_max of defaultNamelists (namelists.F90:135): the default-real literal
# 3.e-5, i.e. 2.9999999242e-05 in double precision
DT_MAX_DEFAULT = float(np.float32(3e-5))

# fdns_params.solver
SOLVERS = {"auto": 0, "lu": 1, "inverse": 2}
# kernel classes of fdns_ens_step_timed, in FDNS_KC_* order (include/fdns.h)
KERNEL_CLASSES = ("nonlin", "imex_pre", "imex_mid", "imex_post", "rk_pre", "rk_mid", "rk_post", "solve_T", "solve_F",
                  "other")


class Tuning(C.Structure):
    """fdns_tuning: launch choices of one (Nr_max, Nm_max); results do not
    depend on them.  tile_waves / tile_stage: one entry per column-tile
    kernel (TILE_KERNELS order)."""
    _fields_ = [("nl_tpb", C.c_int32), ("tile_waves", C.c_int32 * 6), ("tile_stage", C.c_int32 * 6),
                ("tile_pf", C.c_int32 * 6), ("solve_chunk", C.c_int32), ("lane_split", C.c_int32)]

    def as_dict(self):
        return {"nl_tpb": int(self.nl_tpb), "tile_waves": [int(x) for x in self.tile_waves],
                "tile_stage": [int(x) for x in self.tile_stage], "tile_pf": [int(x) for x in self.tile_pf],
                "solve_chunk": int(self.solve_chunk), "lane_split": int(self.lane_split)}


TILE_KERNELS = ("imex_pre", "imex_mid", "imex_post", "rk_pre", "rk_mid", "rk_post")
TILE_OPTIONS = tuple((8 if o & 1 else 4, (o >> 1) & 1, (o >> 2) & 1) for o in range(8))
# fdns_ens_autotune's report: per-step milliseconds of each option tried
TUNE_REPORT = (("nonlin/nl_tpb=128", "nonlin/nl_tpb=256")
               + tuple(f"{k}/waves={w},stage={s},pf={p}" for k in TILE_KERNELS for w, s, p in TILE_OPTIONS)
               + ("solves/chunk=largest", "solves/chunk=16", "solves/chunk=8",
                  "step/lane_split=0", "step/lane_split=1", "step/previous", "step/found"))


# Short names used by the benchmark configurations -> timeschemes.F90 names
SCHEMES = {
    "CNAB2": ("IMEX", "CN", "AB2"),
    "CNAB3": ("IMEX", "CN", "AB3"),
    "SBDF2": ("IMEX", "BDF2", "AB2"),
    "SBDF3": ("IMEX", "BDF3", "AB3"),
    "SBDF4": ("IMEX", "BDF4", "AB4"),
    "ARS443": ("IMEXRK", "ARS443", "expARS443"),
    "BPR353": ("IMEXRK", "BPR353", "expBPR353"),
    "ARS222": ("IMEXRK", "ARS222", "expARS222"),
    "PC2": ("IMEXRK", "PC2", "expPC2"),
    "SSP332": ("IMEXRK", "SSP332", "expSSP332"),
    "SSP543": ("IMEXRK", "SSP543", "expSSP543"),
    "FW53": ("IMEXRK", "FW53", "expFW53"),
    "CK222": ("IMEXRK", "CK222", "expCK222"),
    "BPR442": ("IMEXRK", "BPR442", "expBPR442"),
}
# the IMEXRK schemes whose steps end with Assembly_stage (not stiffly accurate)
ASSEMBLY_SCHEMES = ("EUL111", "ARS232", "ARS233", "ARS343", "MARS343", "IJ3", "PR433", "DBM453", "KC343", "KC564",
                    "KC774", "KC885", "CFN343", "BHR553", "CFN564")
for _n in ASSEMBLY_SCHEMES:
    SCHEMES[_n] = ("IMEXRK", _n, "exp" + _n)
# the IMEXRK schemes whose stages are solved with their own diagonal (LU
# solves).  LZ3A4a is in the registry but refused at creation: its stages 2
# and 4 have aD(k,k) = 0, which makes the vorticity-streamfunction operator
# singular
OWN_DIAGONAL_SCHEMES = ("SMR333", "CB3e", "CB3f", "CB4", "LZ3L1", "LZ4A1")
for _n in OWN_DIAGONAL_SCHEMES:
    SCHEMES[_n] = ("IMEXRK", _n, "exp" + _n)
# the explicit RK path (time_scheme_type = 'RK', time_scheme_imp = 'NONE')
SCHEMES["RK2"] = ("RK", "NONE", "RK2")
SCHEMES["RK4"] = ("RK", "NONE", "RK4")

EXPORTS = ["fdns_create", "fdns_destroy", "fdns_set_state", "fdns_get_state", "fdns_seed_state",
           "fdns_init_perturbation", "fdns_step",
           "fdns_diag", "fdns_get_series", "fdns_budget", "fdns_budget_noalloc", "fdns_spectra", "fdns_refine_state", "fdns_ens_create", "fdns_ens_step", "fdns_ens_destroy",
           "fdns_pool_create", "fdns_pool_submit", "fdns_pool_step", "fdns_pool_run", "fdns_pool_poll", "fdns_pool_result",
           "fdns_pool_release", "fdns_pool_destroy", "fdns_job_footprint", "fdns_strerror", "fdns_device_copy", "fdns_ens_step_timed",
           "fdns_tuning_get", "fdns_tuning_set", "fdns_ens_autotune",
           "fdns_scheme_orders", "fdns_restart", "fdns_restart_dt", "fdns_get_dt", "fdns_get_history", "fdns_write_checkpoint", "fdns_write_snapshot"]


def lib():
    """Load libfdns.so (built by fdns.build); raises if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FdnsError(f"{LIB_PATH} not built: run falcon-dns_amd/fdns/build.py")
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    L.fdns_create.argtypes = [C.POINTER(Params), C.POINTER(vp)]
    L.fdns_destroy.argtypes = [vp]
    L.fdns_set_state.argtypes = [vp, vp, vp, vp, vp]
    L.fdns_get_state.argtypes = [vp, vp, vp, vp, vp, vp]
    L.fdns_seed_state.argtypes = [vp, C.c_uint64, C.c_double]
    L.fdns_init_perturbation.argtypes = [vp, C.c_int, C.c_double]
    L.fdns_restart_dt.argtypes = [vp, C.c_int, vp]
    L.fdns_device_copy.argtypes = [C.c_size_t, C.c_int, C.POINTER(C.c_double)]
    L.fdns_step.argtypes = [vp, C.c_int]
    L.fdns_diag.argtypes = [vp, vp]
    L.fdns_budget.argtypes = [vp, vp, vp]
    L.fdns_spectra.argtypes = [vp, vp, vp]
    L.fdns_refine_state.argtypes = [C.c_int, C.c_int, vp, vp, vp, vp, C.c_int, C.c_int, vp, vp, vp, vp]
    L.fdns_get_series.argtypes = [vp, vp, C.c_int, C.POINTER(C.c_int)]
    L.fdns_ens_create.argtypes = [C.POINTER(vp), C.c_int, C.POINTER(vp)]
    L.fdns_ens_step.argtypes = [vp, C.c_int, C.c_int]
    L.fdns_ens_destroy.argtypes = [vp]
    L.fdns_pool_create.argtypes = [C.c_size_t, C.POINTER(vp)]
    L.fdns_pool_submit.argtypes = [vp, C.POINTER(Params), C.c_uint64, C.c_double, C.c_int, C.POINTER(C.c_int)]
    L.fdns_pool_step.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.fdns_pool_run.argtypes = [vp, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.fdns_pool_poll.argtypes = [vp, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.fdns_pool_result.argtypes = [vp, C.c_int, vp, vp, vp, vp, vp]
    L.fdns_pool_release.argtypes = [vp, C.c_int]
    L.fdns_pool_destroy.argtypes = [vp]
    L.fdns_job_footprint.argtypes = [C.POINTER(Params)]
    L.fdns_job_footprint.restype = C.c_size_t
    L.fdns_strerror.argtypes = [C.c_int]
    L.fdns_strerror.restype = C.c_char_p
    L.fdns_ens_step_timed.argtypes = [vp, C.c_int, vp, vp, vp]
    L.fdns_tuning_get.argtypes = [C.c_int, C.c_int, C.POINTER(Tuning)]
    L.fdns_tuning_set.argtypes = [C.c_int, C.c_int, C.POINTER(Tuning)]
    L.fdns_ens_autotune.argtypes = [vp, C.c_int, C.POINTER(Tuning), vp]
    L.fdns_get_dt.argtypes = [vp, C.POINTER(C.c_double), vp]
    L.fdns_scheme_orders.argtypes = [C.POINTER(Params), C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.fdns_restart.argtypes = [vp, C.c_int, vp, vp, vp, vp, C.c_double]
    L.fdns_get_history.argtypes = [vp] * 7
    L.fdns_write_checkpoint.argtypes = [vp, C.c_char_p]
    L.fdns_write_snapshot.argtypes = [vp, C.c_char_p]
    _lib = L
    return L


def _check(rc):
    if rc != 0:
        msg = lib().fdns_strerror(rc).decode()
        raise FdnsError(f"fdns error {rc}: {msg}")


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def make_params(Nr, Nm, Ra, Pr, dt, scheme="CNAB2", eta=0.35, mBC="NS", buo="IMP", restart=False, n_KE=0,
                solver="auto", vardt=False, CFL=0.9, dt_coef=2.0, dt_max=DT_MAX_DEFAULT, optimizedt=False, lagpts=14):
    """fdns_params; the variable-dt defaults are the reference's
    (namelists.F90:117-140)."""
    stype, imp, exp = SCHEMES[scheme] if isinstance(scheme, str) else scheme
    return Params(Nr=Nr, Nm=Nm, eta=eta, Ra=Ra, Pr=Pr, dt=dt, mBC=mBC.encode(), buo_tscheme=buo.encode(),
                  time_scheme_type=stype.encode(), time_scheme_imp=imp.encode(), time_scheme_exp=exp.encode(),
                  l_restart=int(bool(restart)), n_KE=int(n_KE),
                  solver=SOLVERS[solver] if isinstance(solver, str) else int(solver),
                  l_vartimestep=int(bool(vardt)), l_optimizedt=int(bool(optimizedt)), CFL=CFL, dt_coef=dt_coef,
                  dt_max=dt_max, lagpts=int(lagpts))


def _field(a, shape):
    a = np.asfortranarray(np.asarray(a, dtype=np.complex128))
    if a.shape != shape:
        raise ValueError(f"field shape {a.shape} != {shape}")
    return a


def refine_state(state, Nm_new, Nr_new):
    """The restart's change of resolution (init.F90:722-861) of the four
    fields (tFR, omgFR, urFR, upFR) of an (Nm+1, Nr) checkpoint."""
    Nm_o, Nr_o = state[0].shape[0] - 1, state[0].shape[1]
    ins = [np.asfortranarray(a, dtype=np.complex128) for a in state]
    outs = [np.zeros((Nm_new + 1, Nr_new), dtype=np.complex128, order="F") for _ in range(4)]
    _check(lib().fdns_refine_state(Nm_o, Nr_o, *[_p(a) for a in ins], Nm_new, Nr_new, *[_p(a) for a in outs]))
    return tuple(outs)


class Run:
    """One simulation on the GPU (fdns_create ... fdns_destroy)."""

    def __init__(self, Nr, Nm, Ra, Pr, dt, scheme="CNAB2", eta=0.35, mBC="NS", buo="IMP", restart=False, n_KE=0,
                 solver="auto", **vardt):
        self.params = make_params(Nr, Nm, Ra, Pr, dt, scheme, eta, mBC, buo, restart, n_KE, solver, **vardt)
        self.Nr, self.Nm = Nr, Nm
        h = C.c_void_p()
        _check(lib().fdns_create(C.byref(self.params), C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().fdns_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def shape(self):
        return (self.Nm + 1, self.Nr)

    def set_state(self, tFR, omgFR, urFR, upFR):
        arrs = [_field(a, self.shape) for a in (tFR, omgFR, urFR, upFR)]
        _check(lib().fdns_set_state(self.h, *[_p(a) for a in arrs]))

    def seed(self, seed, amp=1e-3):
        _check(lib().fdns_seed_state(self.h, C.c_uint64(seed), amp))

    def init_perturbation(self, n_init, ampT):
        """The reference's fresh start (init_perturbation, init.F90:277-399)."""
        _check(lib().fdns_init_perturbation(self.h, int(n_init), float(ampT)))

    def get_state(self):
        arrs = [np.zeros(self.shape, dtype=np.complex128, order="F") for _ in range(5)]
        _check(lib().fdns_get_state(self.h, *[_p(a) for a in arrs]))
        return dict(zip(("tFR", "omgFR", "urFR", "upFR", "psi"), arrs))

    def step(self, n=1):
        _check(lib().fdns_step(self.h, int(n)))

    def dt(self, array=False):
        """dt of the last step, or (dt, dt_array) with array=True."""
        v = C.c_double(0.0)
        a = np.zeros(max(self.orders()))
        _check(lib().fdns_get_dt(self.h, C.byref(v), _p(a)))
        return (v.value, a) if array else v.value

    def diag(self):
        out = np.zeros(6)
        _check(lib().fdns_diag(self.h, _p(out)))
        return dict(zip(("t", "KE", "Nu_t", "Nu_b", "KE_r", "KE_phi"), out))

    def budget(self):
        """{KE_mean, vis, buo, t} and the per-level mean-flow balance (Nr x 3:
        du/dt, advection, viscous term) of writeKE_spectral; needs n_KE > 0."""
        out = np.zeros(4)
        ub = np.zeros(3 * self.Nr)
        _check(lib().fdns_budget(self.h, _p(out), _p(ub)))
        return dict(zip(("KE_mean", "vis", "buo", "t"), out)), ub.reshape(self.Nr, 3)

    def spectra(self):
        """calculate_spectra: KEspec(m), Tspec(m), m = 0..Nm."""
        k = np.zeros(self.Nm + 1)
        t = np.zeros(self.Nm + 1)
        _check(lib().fdns_spectra(self.h, _p(k), _p(t)))
        return k, t

    def series(self):
        n = C.c_int(0)
        _check(lib().fdns_get_series(self.h, None, 0, C.byref(n)))
        out = np.zeros((max(n.value, 1), 4))
        _check(lib().fdns_get_series(self.h, _p(out), n.value, C.byref(n)))
        return out[: n.value]

    def orders(self):
        """(n_order_tscheme_imp, n_order_tscheme_exp) of the run's scheme."""
        ni, ne = C.c_int(0), C.c_int(0)
        _check(lib().fdns_scheme_orders(C.byref(self.params), C.byref(ni), C.byref(ne)))
        return ni.value, ne.value

    def restart(self, states, time):
        """Restart start-up from checkpoint states, oldest first: a list of
        (tFR, omgFR, urFR, upFR) tuples; time = the newest one's tot_time."""
        packed = [np.concatenate([_field(st[f], self.shape).ravel(order="F") for st in states])
                  for f in range(4)]
        _check(lib().fdns_restart(self.h, len(states), *[_p(a) for a in packed], float(time)))

    def restart_dt(self, dt_array):
        """dt_array restored from the checkpoints, newest first (after restart)."""
        a = np.ascontiguousarray(dt_array, dtype=np.float64)
        _check(lib().fdns_restart_dt(self.h, len(a), _p(a)))

    def history(self):
        """The checkpoint history arrays: rhs_{imp,exp}_{temp,vort} of shape
        (n, Nm+1, Nr) and rhs_{imp,exp}_uphi_bar (n, Nr), Fortran order."""
        ni, ne = self.orders()
        if self.params.time_scheme_type == b"IMEXRK":
            ni = ne = 1
        Nm1, N = self.Nm + 1, self.Nr
        shapes = [(ni, Nm1, N), (ne, Nm1, N), (ni, Nm1, N), (ne, Nm1, N), (ni, N), (ne, N)]
        arrs = [np.zeros(sh, dtype=np.complex128, order="F") for sh in shapes]
        _check(lib().fdns_get_history(self.h, *[_p(a) for a in arrs]))
        return dict(zip(("imp_temp", "exp_temp", "imp_vort", "exp_vort", "imp_ub", "exp_ub"), arrs))

    def write_checkpoint(self, path):
        _check(lib().fdns_write_checkpoint(self.h, os.fsencode(path)))

    def write_snapshot(self, path):
        _check(lib().fdns_write_snapshot(self.h, os.fsencode(path)))


class Ensemble:
    """Many runs advanced together by one batched step (fdns_ens_*)."""

    def __init__(self, runs):
        self.runs = list(runs)
        arr = (C.c_void_p * len(self.runs))(*[r.h for r in self.runs])
        h = C.c_void_p()
        _check(lib().fdns_ens_create(arr, len(self.runs), C.byref(h)))
        self.h = h

    def step(self, n=1, graph=True):
        _check(lib().fdns_ens_step(self.h, int(n), int(bool(graph))))

    def step_timed(self, n=1):
        """n batched steps without a graph and with HIP events around every
        kernel launch.  Returns (wall_ms, {class: (kernel_ms, launches)})."""
        ms = np.zeros(len(KERNEL_CLASSES))
        cnt = np.zeros(len(KERNEL_CLASSES), dtype=np.int32)
        wall = np.zeros(1)
        _check(lib().fdns_ens_step_timed(self.h, int(n), _p(ms), _p(cnt), _p(wall)))
        return float(wall[0]), {name: (float(ms[c]), int(cnt[c])) for c, name in enumerate(KERNEL_CLASSES)}

    def autotune(self, steps=5):
        """fdns_ens_autotune: returns (chosen tuning dict, {option: ms per
        step} for the options tried)."""
        t = Tuning()
        rep = np.zeros(len(TUNE_REPORT))
        _check(lib().fdns_ens_autotune(self.h, int(steps), C.byref(t), _p(rep)))
        return t.as_dict(), {k: float(v) for k, v in zip(TUNE_REPORT, rep) if v >= 0}

    def close(self):
        if getattr(self, "h", None):
            _check(lib().fdns_ens_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Pool:
    """Serving pool: continuous batching of jobs under an HBM budget."""

    def __init__(self, budget_bytes):
        h = C.c_void_p()
        _check(lib().fdns_pool_create(C.c_size_t(int(budget_bytes)), C.byref(h)))
        self.h = h
        self._shapes = {}

    def submit(self, Nr, Nm, Ra, Pr, dt, scheme, nsteps, seed, amp=1e-3, **kw):
        p = make_params(Nr, Nm, Ra, Pr, dt, scheme, **kw)
        jid = C.c_int(0)
        _check(lib().fdns_pool_submit(self.h, C.byref(p), C.c_uint64(seed), amp, int(nsteps), C.byref(jid)))
        self._shapes[jid.value] = (Nm + 1, Nr)
        return jid.value

    def step(self):
        a = C.c_int(0)
        w = C.c_int(0)
        _check(lib().fdns_pool_step(self.h, C.byref(a), C.byref(w)))
        return a.value, w.value

    def run(self, max_steps):
        """Up to max_steps batched steps in one call (fdns_pool_run):
        returns (steps taken, resident jobs, waiting jobs)."""
        k = C.c_int(0)
        a = C.c_int(0)
        w = C.c_int(0)
        _check(lib().fdns_pool_run(self.h, int(max_steps), C.byref(k), C.byref(a), C.byref(w)))
        return k.value, a.value, w.value

    def poll(self, jid):
        s = C.c_int(0)
        d = C.c_int(0)
        _check(lib().fdns_pool_poll(self.h, int(jid), C.byref(s), C.byref(d)))
        return s.value, d.value

    def result(self, jid):
        shape = self._shapes[jid]
        arrs = [np.zeros(shape, dtype=np.complex128, order="F") for _ in range(4)]
        diag = np.zeros(6)
        _check(lib().fdns_pool_result(self.h, int(jid), *[_p(a) for a in arrs], _p(diag)))
        out = dict(zip(("tFR", "omgFR", "urFR", "upFR"), arrs))
        out["diag"] = dict(zip(("t", "KE", "Nu_t", "Nu_b", "KE_r", "KE_phi"), diag))
        return out

    def release(self, jid):
        _check(lib().fdns_pool_release(self.h, int(jid)))

    def close(self):
        if getattr(self, "h", None):
            lib().fdns_pool_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def get_tuning(Nr, Nm):
    t = Tuning()
    _check(lib().fdns_tuning_get(int(Nr), int(Nm), C.byref(t)))
    return t.as_dict()


def make_tuning(nl_tpb=256, tile_waves=8, tile_stage=0, solve_chunk=0, tile_pf=0, lane_split=0):
    """A Tuning; tile_waves / tile_stage / tile_pf: one value for every
    column-tile kernel or a sequence of six."""
    per = lambda v: list(v) if hasattr(v, "__len__") else [v] * 6
    t = Tuning()
    t.nl_tpb = nl_tpb
    t.tile_waves[:] = per(tile_waves)
    t.tile_stage[:] = per(tile_stage)
    t.tile_pf[:] = per(tile_pf)
    t.solve_chunk = solve_chunk
    t.lane_split = lane_split
    return t


def set_tuning(Nr, Nm, nl_tpb=256, tile_waves=8, tile_stage=0, solve_chunk=0, tile_pf=0, lane_split=0):
    t = make_tuning(nl_tpb, tile_waves, tile_stage, solve_chunk, tile_pf, lane_split)
    _check(lib().fdns_tuning_set(int(Nr), int(Nm), C.byref(t)))


def device_copy(nbytes, reps=5):
    """Mean duration (ms) of a plain device copy of nbytes (k_copy)."""
    ms = C.c_double()
    _check(lib().fdns_device_copy(nbytes, reps, C.byref(ms)))
    return ms.value


def footprint(Nr, Nm, scheme="CNAB2", **kw):
    p = make_params(Nr, Nm, 1e6, 1.0, 1e-6, scheme, **kw)
    return int(lib().fdns_job_footprint(C.byref(p)))
