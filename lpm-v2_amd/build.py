"""Build the gfx950 direct-sum library (liblpm_direct.so) and the Fortran
ISO_C_BINDING smoke program in-tree.

    python lpm-v2_amd/build.py            # library + Fortran program
    python lpm-v2_amd/build.py --lib      # library only

The library is one hipcc link of the csrc/*.hip translation units.  Packed
FP32 instructions are disabled for every device pass.
"""
from __future__ import annotations

import argparse
import os
import shutil
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
FSRC = os.path.join(HERE, "fortran")
LIB = os.path.join(HERE, "liblpm_direct.so")
OBJDIR = os.path.join(HERE, "_build")
FORTRAN_EXE = os.path.join(HERE, "fortran", "lpm_fortran_smoke")

HIP_SOURCES = ["lpm_pair.hip", "lpm_api.hip", "lpm_solver.hip", "lpm_batch.hip", "lpm_mlp.hip", "lpm_ff.hip",
               "lpm_pse_ops.hip"]
EXTRA_SOURCES = []  # appended by later subsystems

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-fPIC",
    "-std=c++17",
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("command failed: " + " ".join(cmd))
    return r


def build_library(verbose: bool = False) -> str:
    os.makedirs(OBJDIR, exist_ok=True)
    hipcc = _hipcc()
    objs = []
    sources = [s for s in HIP_SOURCES + EXTRA_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    procs = []
    for src in sources:
        obj = os.path.join(OBJDIR, src.replace(".hip", ".o"))
        objs.append(obj)
        cmd = [hipcc, *HIPCC_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd))
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)))
    for cmd, p in procs:
        out, err = p.communicate()
        if p.returncode != 0:
            sys.stderr.write(out + err)
            raise RuntimeError("hipcc failed: " + " ".join(cmd))
    tmp = LIB + ".tmp"
    _run([hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    return LIB


def build_fortran(verbose: bool = False) -> str | None:
    """Compile the ISO_C_BINDING module and smoke program with amdflang."""
    flang = shutil.which("amdflang") or "/opt/rocm/bin/amdflang"
    if not os.path.exists(flang):
        return None
    mods = os.path.join(OBJDIR, "fmod")
    os.makedirs(mods, exist_ok=True)
    srcs = [os.path.join(FSRC, f) for f in ("lpm_direct_f.f90", "lpm_bve_solver_f.f90", "lpm_fortran_smoke.f90")]
    srcs = [s for s in srcs if os.path.exists(s)]
    if not srcs:
        return None
    cmd = [flang, "-O2", "-J", mods, "-o", FORTRAN_EXE, *srcs, "-L" + HERE, "-llpm_direct", "-Wl,-rpath," + HERE]
    if verbose:
        print(" ".join(cmd))
    _run(cmd)
    return FORTRAN_EXE


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--lib", action="store_true", help="library only")
    ap.add_argument("-v", "--verbose", action="store_true")
    a = ap.parse_args()
    t = time.time()
    build_library(a.verbose)
    print(f"built {LIB} in {time.time() - t:.1f}s")
    if not a.lib:
        exe = build_fortran(a.verbose)
        if exe:
            print(f"built {exe}")


if __name__ == "__main__":
    main()
