"""Regenerate the golden vectors in tests/golden/ from lpm-v2's own Fortran.

    python tests/golden/make_golden.py

Builds the reference's direct-sum modules and the driver ref_golden.f90 with
oracle/ref.mk (amdflang + the image's MPICH), runs the driver on one MPI rank
in a scratch directory under oracle/_ref/ that links the reference's mesh seed
files, and packs its raw FP64 outputs into tests/golden/ref_<case>.npz (plain
arrays, no pickled objects).  Only this container has /root/reference; the
committed .npz files are what the tests read.
"""
import glob
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = "/root/reference"
OUT = os.path.join(ROOT, "oracle", "_ref")
CASES = ("icos2", "icos3", "cubed3", "quad3")
SEEDS = ("icosTriSeed.dat", "cubedSphereSeed.dat", "quadRectSeed.dat")


def main():
    subprocess.run(["make", "-s", "-f", "oracle/ref.mk"], cwd=ROOT, check=True)
    run = os.path.join(OUT, "run")
    raw = os.path.join(run, "raw")
    os.makedirs(raw, exist_ok=True)
    for f in glob.glob(os.path.join(raw, "*.f64")):
        os.remove(f)
    for s in SEEDS:  # the reference opens its seed files in the working directory
        dst = os.path.join(run, s)
        if not os.path.lexists(dst):
            os.symlink(os.path.join(REF, s), dst)
    r = subprocess.run([os.path.join(OUT, "ref_golden"), raw], cwd=run, capture_output=True, text=True,
                       timeout=600)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise SystemExit(f"ref_golden failed with status {r.returncode}")
    for case in CASES:
        arrays = {}
        for f in sorted(glob.glob(os.path.join(raw, case + "_*.f64"))):
            name = os.path.basename(f)[len(case) + 1:-4]
            arrays[name] = np.fromfile(f, dtype="<f8")
        n = int(arrays.pop("n")[0])
        for k, a in arrays.items():
            assert a.shape == (n,) or a.shape == (1,), (case, k, a.shape)
        dst = os.path.join(HERE, f"ref_{case}.npz")
        np.savez_compressed(dst, n=np.int64(n), **arrays)
        print(f"{dst}: {n} particles, {len(arrays)} arrays")


if __name__ == "__main__":
    main()
