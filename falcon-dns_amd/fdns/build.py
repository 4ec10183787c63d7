"""Build the gfx950 shared library libfdns.so (and the Fortran driver).

Run as a script or call build().  Object files go to falcon-dns_amd/_build,
the library to falcon-dns_amd/lib/libfdns.so, the driver to
falcon-dns_amd/bin/fdns_driver.  Everything compiles here without a GPU.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))   # falcon-dns_amd/
REPO = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
FSRC = os.path.join(PKG, "fortran")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "lib", "libfdns.so")
DRIVER = os.path.join(PKG, "bin", "fdns_driver")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FC = os.environ.get("FDNS_FC", "/opt/rocm/lib/llvm/bin/flang")
ARCH = "gfx950"
# -packed-fp32-ops off: no v_pk_*_f32 from the compiler.  Host code and the
# assembly kernel must not reorder floating-point expressions.
HIPFLAGS = ["--offload-arch=" + ARCH, "-O3", "-fPIC", "-std=c++17", "-Wno-unused-value", "-Wno-unused-result",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
HOSTFLAGS = ["-ffp-contract=off"]

SOURCES = ["fdns_kernels.hip", "fdns_host.cpp"]


def _run(cmd, quiet=False):
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        sys.stderr.write(p.stdout + p.stderr)
        raise RuntimeError("build step failed: " + " ".join(cmd))
    return p


def build_library(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(os.path.dirname(LIB), exist_ok=True)
    objs = []
    for src in SOURCES:
        path = os.path.join(CSRC, src)
        obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
        flags = list(HIPFLAGS)
        if src.endswith(".cpp"):
            flags += HOSTFLAGS
        cmd = [HIPCC] + flags + ["-I", os.path.join(REPO, "include"), "-c", path, "-o", obj]
        if verbose:
            print(" ".join(cmd))
        _run(cmd)
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    return LIB


def build_driver(verbose: bool = False) -> str | None:
    """Fortran driver (iso_c_binding) linked against libfdns.so."""
    if not os.path.exists(FC):
        return None
    os.makedirs(os.path.dirname(DRIVER), exist_ok=True)
    mod_dir = os.path.join(BUILD, "fmod")
    os.makedirs(mod_dir, exist_ok=True)
    srcs = [os.path.join(FSRC, f) for f in ("fdns_core.f90", "fdns_driver.f90")]
    if not all(os.path.exists(s) for s in srcs):
        return None
    libdir = os.path.dirname(LIB)
    cmd = [FC, "-O2", "-module-dir", mod_dir] + srcs + ["-L" + libdir, "-lfdns", "-Wl,-rpath,$ORIGIN/../lib",
                                                       "-o", DRIVER]
    if verbose:
        print(" ".join(cmd))
    _run(cmd)
    return DRIVER


def build(verbose: bool = False) -> str:
    lib = build_library(verbose)
    build_driver(verbose)
    return lib


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
