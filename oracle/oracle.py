"""ctypes wrapper of the CPU oracle (oracle/fdns_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product package.

Spectral fields are numpy complex128 arrays of shape (Nm+1, Nr) in Fortran
order -- the reference's (Nm_max+1, Nr_max) module arrays.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "build", "libfdns_oracle.so")
_lib = None


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc; no GPU needed)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


class Params(C.Structure):
    _fields_ = [
        ("Nr", C.c_int32), ("Nm", C.c_int32), ("eta", C.c_double),
        ("Ra", C.c_double), ("Pr", C.c_double), ("dt", C.c_double),
        ("mBC", C.c_int32), ("buo_imp", C.c_int32), ("restart", C.c_int32),
        ("nthreads", C.c_int32),
        ("scheme_type", C.c_char * 16), ("scheme_imp", C.c_char * 16),
        ("scheme_exp", C.c_char * 16),
        ("vardt", C.c_int32), ("optimizedt", C.c_int32),
        ("CFL", C.c_double), ("dt_coef", C.c_double), ("dt_max", C.c_double),
        ("lagpts", C.c_int32),
    ]


_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


def lib():
    global _lib
    if _lib is None:
        build()
        L = C.CDLL(_LIB_PATH)
        vp = C.c_void_p
        L.orc_create.argtypes = [C.POINTER(Params), C.POINTER(vp)]
        L.orc_create.restype = C.c_int
        L.orc_destroy.argtypes = [vp]
        L.orc_set_state.argtypes = [vp, vp, vp, vp, vp]
        L.orc_get_state.argtypes = [vp, vp, vp, vp, vp, vp]
        L.orc_step.argtypes = [vp, C.c_int]
        L.orc_restart.argtypes = [vp, C.c_int, vp, vp, vp, vp, C.c_double]
        L.orc_restart.restype = C.c_int
        L.orc_get_history.argtypes = [vp] * 7
        L.orc_diag.argtypes = [vp, vp]
        L.orc_dt.argtypes = [vp]
        L.orc_dt.restype = C.c_double
        L.orc_get_TFC.argtypes = [vp, vp]
        L.orc_rk_mean_rhs.argtypes = [vp, vp]
        L.orc_budget.argtypes = [vp, vp, vp]
        L.orc_refine.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, vp, vp]
        L.orc_spectra.argtypes = [vp, vp, vp]
        L.orc_seed_state.argtypes = [C.c_int, C.c_int, C.c_double, C.c_uint64, C.c_double, vp, vp, vp, vp]
        L.orc_dcheb.argtypes = [C.c_int, vp, vp, vp, vp]
        L.orc_grid.argtypes = [C.c_int, C.c_double, vp, vp, vp]
        L.orc_fft.argtypes = [C.c_int, vp, vp, C.c_int]
        for f in ("orc_chebtransform", "orc_chebinvtran", "orc_chebinvtranD1", "orc_chebinvtranD2"):
            getattr(L, f).argtypes = [C.c_int, vp, vp]
        L.orc_forfft.argtypes = [C.c_int, C.c_int, vp, vp]
        L.orc_invfft.argtypes = [C.c_int, C.c_int, vp, vp]
        L.orc_radint.argtypes = [C.c_int, C.c_double, vp]
        L.orc_radint.restype = C.c_double
        L.orc_mat_build.argtypes = [C.c_int, C.c_double, C.c_int, C.c_double, C.c_double, C.c_int, C.c_int,
                                    vp, vp, vp]
        L.orc_mat_build_lap.argtypes = [C.c_int, C.c_double, C.c_int, vp]
        L.orc_dgetrf.argtypes = [C.c_int, vp, vp]
        L.orc_dgetrf.restype = C.c_int
        L.orc_dgetrs.argtypes = [C.c_int, vp, vp, vp]
        L.orc_nonlin.argtypes = [C.c_int, C.c_int] + [vp] * 9
        L.orc_scheme_info.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p] + [vp] * 7
        L.orc_scheme_info.restype = C.c_int
        L.orc_perturbation_field.argtypes = [C.c_int, C.c_int, C.c_double, C.c_int, C.c_double, vp]
        L.orc_init_perturbation.argtypes = [C.c_int, C.c_int, C.c_double, C.c_int, C.c_double, vp, vp, vp, vp]
        L.orc_gaspari.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double]
        L.orc_gaspari.restype = C.c_double
        L.orc_restart_dt.argtypes = [vp, C.c_int, vp]
        L.orc_restart_dt.restype = C.c_int
        L.orc_set_mutation.argtypes = [C.c_int]
        L.orc_get_mutation.restype = C.c_int
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def _cf(a, shape=None):
    """complex128, contiguous in the reference's column-major layout."""
    a = np.asarray(a, dtype=np.complex128)
    if shape is not None:
        assert a.shape == shape, (a.shape, shape)
    return np.asfortranarray(a)


# ---------------------------------------------------------------- blocks
def dcheb(N):
    t = np.zeros((N, N), order="F"); D = np.zeros((N, N), order="F"); D2 = np.zeros((N, N), order="F")
    xp = np.zeros(N)
    lib().orc_dcheb(N, _p(t), _p(D), _p(D2), _p(xp))
    return t, D, D2, xp


def grid(N, eta):
    r = np.zeros(N); rr = np.zeros(N); rr2 = np.zeros(N)
    lib().orc_grid(N, eta, _p(r), _p(rr), _p(rr2))
    return r, rr, rr2


def fft(x, sign):
    x = np.ascontiguousarray(x, dtype=np.complex128)
    y = np.zeros_like(x)
    lib().orc_fft(len(x), _p(x), _p(y), sign)
    return y


def _cheb1(fname, v):
    v = np.ascontiguousarray(v, dtype=np.complex128)
    out = np.zeros_like(v)
    getattr(lib(), fname)(len(v), _p(v), _p(out))
    return out


def chebtransform(f):
    return _cheb1("orc_chebtransform", f)


def chebinvtran(ft):
    return _cheb1("orc_chebinvtran", ft)


def chebinvtranD1(ft):
    return _cheb1("orc_chebinvtranD1", ft)


def chebinvtranD2(ft):
    return _cheb1("orc_chebinvtranD2", ft)


def forfft(x, Nm):
    x = np.ascontiguousarray(x, dtype=np.float64)
    f = np.zeros(Nm + 1, dtype=np.complex128)
    lib().orc_forfft(Nm, len(x), _p(x), _p(f))
    return f


def invfft(f, Np):
    f = np.ascontiguousarray(f, dtype=np.complex128)
    x = np.zeros(Np)
    lib().orc_invfft(len(f) - 1, Np, _p(f), _p(x))
    return x


def radint(f, eta):
    f = np.ascontiguousarray(f, dtype=np.float64)
    return lib().orc_radint(len(f), eta, _p(f))


def mat_build(N, eta, n, wt_dt, Pr, mBC="NS", irk_order=False):
    AT = np.zeros((N, N), order="F"); AF = np.zeros((2 * N, 2 * N), order="F"); Au = np.zeros((N, N), order="F")
    lib().orc_mat_build(N, eta, n, wt_dt, Pr, 0 if mBC == "NS" else 1, int(irk_order), _p(AT), _p(AF), _p(Au))
    return AT, AF, Au


def dgetrf(A):
    A = np.asfortranarray(np.array(A, dtype=np.float64))
    n = A.shape[0]
    piv = np.zeros(n, dtype=np.int32)
    info = lib().orc_dgetrf(n, _p(A), _p(piv))
    return A, piv, info


def dgetrs(LU, piv, b):
    b = np.array(b, dtype=np.float64, copy=True)
    lib().orc_dgetrs(LU.shape[0], _p(np.asfortranarray(LU)), _p(np.ascontiguousarray(piv, dtype=np.int32)), _p(b))
    return b


def nonlin(tFR, omgFR, upFR, urFR):
    Nm1, Nr = tFR.shape
    ins = [_cf(a) for a in (tFR, omgFR, upFR, urFR)]
    outs = [np.zeros((Nm1, Nr), dtype=np.complex128, order="F") for _ in range(4)]
    mean = np.zeros(Nr)
    lib().orc_nonlin(Nm1 - 1, Nr, *[_p(a) for a in ins], *[_p(a) for a in outs], _p(mean))
    return outs + [mean]


def scheme_info(stype, imp, exp):
    ni = np.zeros(1, dtype=np.int32); ne = np.zeros(1, dtype=np.int32); sa = np.zeros(1, dtype=np.int32)
    aD = np.zeros(64); aA = np.zeros(64); bD = np.zeros(8); bA = np.zeros(8)
    rc = lib().orc_scheme_info(stype.encode(), imp.encode(), exp.encode(), _p(ni), _p(ne),
                               _p(aD), _p(bD), _p(aA), _p(bA), _p(sa))
    if rc:
        raise ValueError(f"scheme {stype}/{imp}/{exp} not available (rc={rc})")
    S = int(ne[0])
    out = {"n_imp": int(ni[0]), "n_exp": S}
    if stype in ("IMEXRK", "RK"):
        out.update(aD=aD[:S * S].reshape(S, S).copy(), aA=aA[:S * S].reshape(S, S).copy(),
                   bD=bD[:S].copy(), bA=bA[:S].copy(), sa=bool(sa[0]))
    return out


class mutation:
    """Context manager planting a mistake in the oracle's coupled right-hand
    side (1: buoyancy sign, 2: Ra instead of Ra/Pr, 3: sign of (u_r f)/r in
    the flux-form advection); the tests that pin that right-hand side use it to
    show their checks fail on such a slip.  0 is the restatement."""

    def __init__(self, k):
        self.k = int(k)

    def __enter__(self):
        lib().orc_set_mutation(self.k)
        return self

    def __exit__(self, *exc):
        lib().orc_set_mutation(0)
        return False


def refine(state, Nm_new, Nr_new):
    """The restart's change of resolution (init.F90:722-861) of the four
    fields (tFR, omgFR, urFR, upFR) of an (Nm+1, Nr) checkpoint."""
    Nm_o, Nr_o = state[0].shape[0] - 1, state[0].shape[1]
    ins = [np.asfortranarray(a, dtype=np.complex128) for a in state]
    outs = [np.zeros((Nm_new + 1, Nr_new), dtype=np.complex128, order="F") for _ in range(4)]
    pin = (C.c_void_p * 4)(*[a.ctypes.data for a in ins])
    pout = (C.c_void_p * 4)(*[a.ctypes.data for a in outs])
    rc = lib().orc_refine(Nm_o, Nr_o, Nm_new, Nr_new, C.cast(pin, C.c_void_p), C.cast(pout, C.c_void_p))
    if rc:
        raise ValueError(f"refine {Nm_o}x{Nr_o} -> {Nm_new}x{Nr_new} not handled (rc={rc})")
    return tuple(outs)


def seed_state(Nr, Nm, eta, seed, amp=1e-3):
    arrs = [np.zeros((Nm + 1, Nr), dtype=np.complex128, order="F") for _ in range(4)]
    lib().orc_seed_state(Nr, Nm, eta, seed, amp, *[_p(a) for a in arrs])
    return tuple(arrs)   # tFR, omgFR, urFR, upFR


def perturbation_field(Nr, Nm, eta, n_init, ampT):
    """init_perturbation's physical temperature tt(Np, Nr) (conduction included)."""
    tt = np.zeros((3 * Nm, Nr), order="F")
    lib().orc_perturbation_field(Nr, Nm, eta, int(n_init), float(ampT), _p(tt))
    return tt


def init_perturbation(Nr, Nm, eta, n_init, ampT):
    """init_perturbation's spectral state (tFR, omgFR, urFR, upFR)."""
    arrs = [np.zeros((Nm + 1, Nr), dtype=np.complex128, order="F") for _ in range(4)]
    lib().orc_init_perturbation(Nr, Nm, eta, int(n_init), float(ampT), *[_p(a) for a in arrs])
    return tuple(arrs)


def gaspari(piece, r, c, L):
    """Gaspari-Cohn piece of init.F90:937-985: 'middle', 'edge' or 'center'."""
    return lib().orc_gaspari({"middle": 0, "edge": 1, "center": 2}[piece], float(r), float(c), float(L))


# dt_max of defaultNamelists (namelists.F90:135): the default-real literal
# 3.e-5, i.e. 2.9999999242e-05 in double precision
DT_MAX_DEFAULT = float(np.float32(3e-5))

# ---------------------------------------------------------------- run
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
# the IMEXRK schemes the step ends with Assembly_stage for (not stiffly accurate)
for _n in ("EUL111", "ARS232", "ARS233", "ARS343", "MARS343", "IJ3", "PR433", "DBM453", "KC343", "KC564",
           "KC774", "KC885", "CFN343", "BHR553", "CFN564"):
    SCHEMES[_n] = ("IMEXRK", _n, "exp" + _n)
# the IMEXRK schemes whose stages are solved with their own diagonal (diag_diff)
for _n in ("SMR333", "CB3e", "CB3f", "CB4", "LZ3L1", "LZ3A4a", "LZ4A1"):
    SCHEMES[_n] = ("IMEXRK", _n, "exp" + _n)
# the explicit RK path (time_scheme_type = 'RK', time_scheme_imp = 'NONE')
SCHEMES["RK2"] = ("RK", "NONE", "RK2")
SCHEMES["RK4"] = ("RK", "NONE", "RK4")


class Run:
    """One oracle run: timeloop_imex / timeloop_imexrk / timeloop_rk of the reference."""

    def __init__(self, Nr, Nm, Ra, Pr, dt, scheme="CNAB2", eta=0.35, mBC="NS", buo="IMP",
                 restart=False, nthreads=0, vardt=False, CFL=0.9, dt_coef=2.0, dt_max=DT_MAX_DEFAULT, optimizedt=False,
                 lagpts=14):
        stype, imp, exp = SCHEMES[scheme] if scheme in SCHEMES else scheme
        p = Params(Nr=Nr, Nm=Nm, eta=eta, Ra=Ra, Pr=Pr, dt=dt, mBC=0 if mBC == "NS" else 1,
                   buo_imp=1 if buo == "IMP" else 0, restart=int(bool(restart)), nthreads=nthreads,
                   scheme_type=stype.encode(), scheme_imp=imp.encode(), scheme_exp=exp.encode(),
                   vardt=int(bool(vardt)), optimizedt=int(bool(optimizedt)), CFL=CFL, dt_coef=dt_coef,
                   dt_max=dt_max, lagpts=lagpts)
        h = C.c_void_p()
        rc = lib().orc_create(C.byref(p), C.byref(h))
        if rc:
            raise ValueError(f"orc_create failed rc={rc}")
        self.h = h
        self.Nr, self.Nm, self.eta = Nr, Nm, eta

    def __del__(self):
        h = getattr(self, "h", None)
        if h:
            lib().orc_destroy(h)
            self.h = None

    def set_state(self, tFR, omgFR, urFR, upFR):
        s = (self.Nm + 1, self.Nr)
        self._keep = [_cf(a, s) for a in (tFR, omgFR, urFR, upFR)]
        lib().orc_set_state(self.h, *[_p(a) for a in self._keep])

    def get_state(self):
        s = (self.Nm + 1, self.Nr)
        arrs = [np.zeros(s, dtype=np.complex128, order="F") for _ in range(5)]
        lib().orc_get_state(self.h, *[_p(a) for a in arrs])
        return dict(zip(("tFR", "omgFR", "urFR", "upFR", "psi"), arrs))

    def step(self, n=1):
        lib().orc_step(self.h, int(n))

    def restart(self, states, t):
        """Restart start-up from checkpoint states, oldest first: a list of
        (tFR, omgFR, urFR, upFR) tuples; t = the newest one's tot_time."""
        s = (self.Nm + 1, self.Nr)
        K = len(states)
        packed = [np.concatenate([_cf(st[f], s).ravel(order="F") for st in states]) for f in range(4)]
        self._keep = packed
        rc = lib().orc_restart(self.h, K, *[_p(a) for a in packed], float(t))
        if rc:
            raise ValueError(f"orc_restart refused K={K} (rc={rc})")

    def restart_dt(self, dt_array):
        """dt_array restored from the checkpoints, newest first (after restart)."""
        a = np.ascontiguousarray(dt_array, dtype=np.float64)
        if lib().orc_restart_dt(self.h, len(a), _p(a)):
            raise ValueError("dt_array of 1..8 entries")

    def history(self, n_imp, n_exp):
        """Checkpoint history arrays in the reference's layout: rhs_*_temp,
        rhs_*_vort of shape (n, Nm+1, Nr) and rhs_*_uphi_bar (n, Nr)."""
        Nm1, N = self.Nm + 1, self.Nr
        shapes = [(n_imp, Nm1, N), (n_exp, Nm1, N), (n_imp, Nm1, N), (n_exp, Nm1, N), (n_imp, N), (n_exp, N)]
        arrs = [np.zeros(sh, dtype=np.complex128, order="F") for sh in shapes]
        lib().orc_get_history(self.h, *[_p(a) for a in arrs])
        return dict(zip(("imp_temp", "exp_temp", "imp_vort", "exp_vort", "imp_ub", "exp_ub"), arrs))

    def dt(self):
        """dt_new of the last step."""
        return lib().orc_dt(self.h)

    def diag(self):
        out = np.zeros(6)
        lib().orc_diag(self.h, _p(out))
        return dict(zip(("t", "KE", "Nu_t", "Nu_b", "KE_r", "KE_phi"), out))

    def TFC(self):
        v = np.zeros(self.Nr, dtype=np.complex128)
        lib().orc_get_TFC(self.h, _p(v))
        return v

    def budget(self):
        """{KE_mean, vis, buo, t} and the per-level mean-flow balance
        (Nr x 3: du/dt, advection, viscous term) of writeKE_spectral."""
        out = np.zeros(4)
        ub = np.zeros(3 * self.Nr)
        lib().orc_budget(self.h, _p(out), _p(ub))
        return dict(zip(("KE_mean", "vis", "buo", "t"), out)), ub.reshape(self.Nr, 3)

    def spectra(self):
        """calculate_spectra: KEspec(m), Tspec(m)."""
        k = np.zeros(self.Nm + 1)
        t = np.zeros(self.Nm + 1)
        lib().orc_spectra(self.h, _p(k), _p(t))
        return k, t

    def rk_mean_rhs(self):
        """The RK path's mean-flow right-hand side of the current state."""
        v = np.zeros(self.Nr, dtype=np.complex128)
        if lib().orc_rk_mean_rhs(self.h, _p(v)):
            raise ValueError("not an RK run")
        return v

    def run(self, nsteps, every=1):
        """Advance nsteps, recording (t, KE, Nu_t, Nu_b) every `every` steps."""
        series = []
        done = 0
        while done < nsteps:
            k = min(every, nsteps - done)
            self.step(k)
            done += k
            d = self.diag()
            series.append((d["t"], d["KE"], d["Nu_t"], d["Nu_b"]))
        return np.array(series)
