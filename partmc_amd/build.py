"""Build the gfx950 shared library ``libpartmc_coag.so`` in-tree.

``python -m partmc_amd.build`` (or ``__graft_entry__.build()``) compiles
``csrc/pmc_coag.hip`` with hipcc for gfx950 only.  FP64 device code keeps
every multiply and add separately rounded (``-ffp-contract=off``), matching
the reference's FP64 arithmetic; packed FP32 instructions are disabled.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libpartmc_coag.so")
SOURCES = [os.path.join(CSRC, "pmc_coag.hip")]
DEPS = SOURCES + [os.path.join(CSRC, f) for f in
                  ("pmc_device.h", "pmc_state.h", "pmc_exact.h", "pmc_parallel.h",
                                   "pmc_parallel_step.h", "pmc_stats.h", "pmc_rebalance.h",
                                   "pmc_scan.h")] + [
    os.path.join(HERE, "..", "include", "partmc_coag.h")]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def command(out: str = LIB) -> list:
    return [hipcc(), "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
            "-ffp-contract=off", "-fno-fast-math",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wno-unused-command-line-argument",
            "-o", out] + SOURCES


FORTRAN = os.path.join(HERE, "fortran")
DRIVER = os.path.join(FORTRAN, "coag_driver")


def build_fortran(verbose: bool = False):
    """The Fortran ISO_C_BINDING module and its driver program (amdflang),
    linked against the library with an $ORIGIN-relative runpath."""
    fc = shutil.which("amdflang") or "/opt/rocm/bin/amdflang"
    if not os.path.exists(fc):
        if verbose:
            print("amdflang not found: Fortran driver not built", flush=True)
        return None
    import tempfile
    with tempfile.TemporaryDirectory() as mod_dir:
        cmd = [fc, "-O2", "-J", mod_dir, "-o", DRIVER + ".tmp",
               os.path.join(FORTRAN, "pmc_coag_gpu.F90"), os.path.join(FORTRAN, "coag_driver.F90"),
               "-L" + HERE, "-lpartmc_coag", "-Wl,-rpath,$ORIGIN/.."]
        if verbose:
            print(" ".join(cmd), flush=True)
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            sys.stderr.write(res.stdout + res.stderr)
            raise RuntimeError("amdflang failed building the Fortran driver")
    os.replace(DRIVER + ".tmp", DRIVER)
    return DRIVER


PROBE = os.path.join(HERE, "libpmc_probe.so")


def build_probe(verbose: bool = False) -> str:
    """The throughput probes (csrc/pmc_probe.hip) that bench.py --probe runs."""
    tmp = PROBE + ".tmp"
    cmd = [hipcc(), "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           "-Wno-unused-command-line-argument", "-o", tmp, os.path.join(CSRC, "pmc_probe.hip")]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed building libpmc_probe.so")
    os.replace(tmp, PROBE)
    return PROBE


def build(verbose: bool = False) -> str:
    tmp = LIB + ".tmp"
    cmd = command(tmp)
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed building libpartmc_coag.so")
    os.replace(tmp, LIB)
    build_probe(verbose)
    build_fortran(verbose)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
