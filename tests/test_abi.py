"""The C-ABI library on the CPU: it loads, exports every entry point declared
in include/fdns.h, agrees with the header on the parameter struct layout, and
its host-side validation answers with the documented status codes before any
device call (no GPU is needed for these)."""
from __future__ import annotations

import ctypes as C
import os
import re
import subprocess

import pytest

from conftest import REPO

HEADER = os.path.join(REPO, "include", "fdns.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[\w\s\*]*?\b(fdns_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


@pytest.fixture(scope="module")
def lib():
    import fdns
    if not os.path.exists(fdns.LIB_PATH):
        from fdns import build
        build.build_library()
    return fdns.lib()


def test_header_declares_the_boundary():
    names = declared_functions()
    for must in ("fdns_create", "fdns_destroy", "fdns_set_state", "fdns_get_state", "fdns_step", "fdns_diag",
                 "fdns_ens_create", "fdns_ens_step", "fdns_pool_create", "fdns_pool_submit", "fdns_pool_step",
                 "fdns_pool_poll", "fdns_pool_result", "fdns_strerror"):
        assert must in names


def test_library_exports_every_declared_symbol(lib):
    import fdns
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", fdns.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\bT (fdns_\w+)$", out, flags=re.M))
    assert set(declared_functions()) <= exported
    assert set(fdns.EXPORTS) == set(declared_functions())


def test_params_layout_matches_header(tmp_path):
    import fdns
    src = tmp_path / "layout.c"
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "fdns.h"\n'
                   "int main(void){printf(\"%zu\", sizeof(fdns_params));"
                   + "".join(f'printf(" %zu", offsetof(fdns_params, {f}));' for f, _ in fdns.Params._fields_)
                   + "return 0;}\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.dirname(HEADER), str(src), "-o", str(exe)], check=True)
    vals = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    assert vals[0] == C.sizeof(fdns.Params)
    assert vals[1:] == [getattr(fdns.Params, f).offset for f, _ in fdns.Params._fields_]


def create_rc(lib, **kw):
    import fdns
    args = dict(Nr=33, Nm=16, Ra=1e5, Pr=1.0, dt=1e-5, scheme="CNAB2")
    args.update(kw)
    p = fdns.make_params(**args)
    h = C.c_void_p()
    return lib.fdns_create(C.byref(p), C.byref(h)), h


@pytest.mark.parametrize("kw,code", [
    (dict(scheme=("IMEX", "BDF7", "AB2")), -2),
    (dict(scheme=("IMEXRK", "ARS443", "expBPR353")), -2),
    (dict(scheme=("RK", "RK4", "RK4")), -2),
    (dict(eta=1.5), -1),
    (dict(dt=0.0), -1),
    (dict(Pr=-1.0), -1),
    (dict(mBC="XX"), -1),
    (dict(buo="BOTH"), -1),
    (dict(Nr=600), -3),
    (dict(Nr=3), -3),
    (dict(Nm=2000), -3),
    (dict(solver=3), -1),
])
def test_create_rejects_invalid_parameters(lib, kw, code):
    rc, h = create_rc(lib, **kw)
    assert rc == code
    assert not h.value
    assert lib.fdns_strerror(rc)


def test_strerror_codes(lib):
    for code in range(0, -9, -1):
        assert lib.fdns_strerror(code)


def test_null_arguments(lib):
    h = C.c_void_p()
    assert lib.fdns_create(None, C.byref(h)) == -1
    assert lib.fdns_destroy(None) == 0


def test_job_footprint(lib):
    import fdns
    fp = lambda **kw: lib.fdns_job_footprint(C.byref(fdns.make_params(**{**dict(Nr=65, Nm=128, Ra=1e6, Pr=1.0,
                                                                                    dt=1e-6), **kw})))
    a = fp()
    assert a > 0
    assert fp(Nr=129) > a and fp(Nm=256) > a
    assert fp(scheme="ARS443") > fp(scheme="CNAB2")      # stage arrays exceed two-level histories
    # the resident operator inverses dominate: AT^{-1} (N^2 per mode), the
    # interior-omega columns of AF^{-1} (2N(N-2) per mode >= 1), A_uphi^{-1}
    N, Nm = 65, 128
    ops = 8 * (N * N * (Nm + 1) + 2 * N * (N - 2) * Nm + N * N)
    assert ops <= a < ops + 8 * 64 * N * (Nm + 1) * 16
    assert fp(scheme=("IMEX", "BDF9", "AB2")) == 0
    # LU mode keeps the full packed factors: AT (N^2) and AF (4N^2) per mode
    lu = fp(solver="lu")
    assert lu > a and lu >= 8 * 5 * N * N * Nm
    assert fp(solver="inverse") == a                    # auto picks the inverses at Nr <= 129
    assert fp(Nr=257, solver="auto") == fp(Nr=257, solver="lu")
    assert fp(solver=7) == 0


def test_tuning_table(lib):
    """fdns_tuning_set / get: host-side table of launch choices (no GPU call)."""
    import fdns
    T = fdns.make_tuning
    t = fdns.Tuning()
    assert lib.fdns_tuning_get(41, 30, C.byref(t)) == 0
    assert t.as_dict() == dict(nl_tpb=256, tile_waves=[8] * 6, tile_stage=[0] * 6, tile_pf=[0] * 6, solve_chunk=0,
                               lane_split=0)
    for bad in (T(64, 8, 0, 0), T(256, 6, 0, 0), T(256, 8, 2, 0), T(256, 8, 0, 12), T(256, 8, 0, 40),
                T(256, 8, 0, -8), T(256, [8, 8, 8, 8, 8, 5], 0, 0), T(256, 8, [0, 0, 0, 0, 0, 3], 0),
                T(256, 8, 0, 0, [0, 2, 0, 0, 0, 0]), T(256, 8, 0, 0, 0, 2)):
        assert lib.fdns_tuning_set(41, 30, C.byref(bad)) == -1
    assert lib.fdns_tuning_set(41, 30, None) == -1
    assert lib.fdns_tuning_get(41, 30, None) == -1
    good = T(128, [4, 8, 4, 8, 4, 8], [1, 0, 0, 1, 1, 0], 16, [1, 1, 0, 0, 1, 0], 1)
    assert lib.fdns_tuning_set(41, 30, C.byref(good)) == 0
    assert lib.fdns_tuning_get(41, 30, C.byref(t)) == 0 and t.as_dict() == good.as_dict()
    assert lib.fdns_tuning_get(41, 31, C.byref(t)) == 0 and t.nl_tpb == 256
    # the library's table: the benchmark shape starts from its stored choice, a valid one
    assert lib.fdns_tuning_get(97, 256, C.byref(t)) == 0
    assert lib.fdns_tuning_set(97, 256, C.byref(t)) == 0
    # and so does every served shape (Nr 33..257 x Nm 64..768)
    plain = T().as_dict()
    tuned = 0
    for nr in (33, 65, 129, 257):
        for nm in (64, 128, 256, 512, 768):
            assert lib.fdns_tuning_get(nr, nm, C.byref(t)) == 0
            assert lib.fdns_tuning_set(nr, nm, C.byref(t)) == 0
            tuned += t.as_dict() != plain
    assert tuned >= 15
    assert lib.fdns_tuning_set(41, 30, C.byref(T())) == 0
    assert lib.fdns_ens_autotune(None, 3, None, None) == -1
