"""Compile the HIP library liboqg.so in-tree for gfx950.

Usage: python -m openqg_amd.build   (or openqg_amd.build.build()).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")
LIB = os.path.join(HERE, "liboqg.so")
SOURCES = ["oqg_stencil.hip", "oqg_transform.hip", "oqg_adjoint.hip", "oqg_forcing.hip", "oqg_diag.hip",
           "oqg_mixed.hip", "oqg_couple.hip", "oqg_api.hip"]
WAVE_SOURCE = "oqg_wave.hip"
WAVE_LAYERS = range(2, 9)   # OQG_MAX_LAYERS
ARCH = os.environ.get("OQG_OFFLOAD_ARCH", "gfx950")

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # FP64 semantics: one IEEE rounding per operation, as in the reference restatement
    "-ffp-contract=off",
    # no packed FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose: bool = False) -> str:
    """Compile every source and link liboqg.so; returns the library path.

    Objects go to a private temporary directory, so an interrupted build or two
    builds at once leave nothing behind in the source tree."""
    objdir = tempfile.mkdtemp(prefix="oqg_build_")
    try:
        return _build_in(objdir, verbose)
    finally:
        shutil.rmtree(objdir, ignore_errors=True)


def _build_in(objdir: str, verbose: bool) -> str:
    objs = []
    procs = []
    units = [(src, src.replace(".hip", ".o"), []) for src in SOURCES]
    # the wavefront-transform kernels, one object per layer count (compiled in parallel)
    units += [(WAVE_SOURCE, f"oqg_wave_nl{nl}.o", [f"-DOQG_WAVE_NL={nl}"]) for nl in WAVE_LAYERS]
    for src, oname, extra in units:
        obj = os.path.join(objdir, oname)
        cmd = [hipcc(), *HIPCC_FLAGS, *extra, f"-I{INCLUDE}", "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        objs.append(obj)
    for cmd, p in procs:
        out, _ = p.communicate()
        text = out.decode(errors="replace")
        if p.returncode != 0:
            raise RuntimeError(f"hipcc failed ({' '.join(cmd)}):\n{text}")
        if verbose and text.strip():
            print(text)
    tmp = os.path.join(objdir, "liboqg.so")
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout.decode(errors='replace')}")
    shutil.copyfile(tmp, LIB + ".tmp")
    os.chmod(LIB + ".tmp", 0o755)
    os.replace(LIB + ".tmp", LIB)
    return LIB


FORTRAN = os.path.join(HERE, "fortran")
FORTRAN_BUILD = os.path.join(FORTRAN, "build")
FORTRAN_TEST = os.path.join(FORTRAN_BUILD, "test_oqg_c")
# the drop-in module core, compiled against a test-only stand-in for OpenQG's
# module qg (tests/fortran) into a program that drives it like driver.F90 does
FORTRAN_CORE_TEST = os.path.join(FORTRAN_BUILD, "test_core")
TESTS_FORTRAN = os.path.join(os.path.dirname(HERE), "tests", "fortran")


def flang() -> str | None:
    for cand in (os.environ.get("FC_OQG"), shutil.which("amdflang"), "/opt/rocm/bin/amdflang"):
        if cand and os.path.exists(cand):
            return cand
    return None


def build_fortran(verbose: bool = False) -> str | None:
    """Compile the ISO_C_BINDING interface module and the Fortran test program
    against liboqg.so (needs amdflang; returns None when it is absent)."""
    fc = flang()
    if fc is None:
        return None
    os.makedirs(FORTRAN_BUILD, exist_ok=True)
    cmd = [fc, "-O2", "-J", FORTRAN_BUILD, os.path.join(FORTRAN, "oqg_c.F90"), os.path.join(FORTRAN, "test_oqg_c.F90"),
           "-L", HERE, "-loqg", "-Wl,-rpath,$ORIGIN/../..", "-o", FORTRAN_TEST]
    if verbose:
        print(" ".join(cmd), flush=True)
    _run_fc(cmd)
    stub = os.path.join(TESTS_FORTRAN, "qg_types.F90")
    if os.path.exists(stub):
        # oqg_c, module qg, then core.F90 (which uses both), then the program
        cmd = [fc, "-O2", "-J", FORTRAN_BUILD, os.path.join(FORTRAN, "oqg_c.F90"), stub,
               os.path.join(FORTRAN, "core.F90"), os.path.join(TESTS_FORTRAN, "test_core.F90"),
               "-L", HERE, "-loqg", "-Wl,-rpath,$ORIGIN/../..", "-o", FORTRAN_CORE_TEST]
        if verbose:
            print(" ".join(cmd), flush=True)
        _run_fc(cmd)
    return FORTRAN_TEST


def _run_fc(cmd):
    r = subprocess.run(cmd, cwd=FORTRAN_BUILD, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    if r.returncode != 0:
        raise RuntimeError(f"amdflang failed:\n{r.stdout.decode(errors='replace')}")


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
    print(build_fortran(verbose="-v" in sys.argv))
