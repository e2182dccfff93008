"""Writes tests/golden/pdgemm_<case>.npz: ScaLAPACK's own pdgemm on a 1 x 1
BLACS grid (MKL ScaLAPACK + MPICH from /opt/conda in the build container).

The reference tests its pdgemm against ScaLAPACK's (tests/pdgemm.cpp:420-905,
to 1e-6); these fixtures carry ScaLAPACK's answers for submatrix offsets
(ia, ja, ib, jb, ic, jc), all transposes, alpha/beta and block sizes, so the
GPU path and the oracle's descriptor handling can be checked against
ScaLAPACK itself.  Inputs are regenerated from their seeds (SplitMix64
uniform over the whole local arrays, see pdgemm_scalapack.c); each fixture
stores the case, the SHA-256 of the regenerated inputs and ScaLAPACK's whole
output array.
"""
import hashlib
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
import oracle  # noqa: E402

CONDA = "/opt/conda"
CASES = [
    # name, ta, tb, m, n, k, alpha, beta, ia, ja, ib, jb, ic, jc, mb, nb, pad
    ("NN_1x1", "N", "N", 37, 29, 41, 1.0, 0.0, 1, 1, 1, 1, 1, 1, 8, 8, 0),
    ("TN_sub", "T", "N", 33, 21, 45, -1.3, 0.5, 3, 2, 2, 4, 5, 3, 16, 8, 2),
    ("NT_sub", "N", "T", 25, 31, 19, 0.75, 1.0, 2, 5, 4, 1, 1, 2, 4, 16, 1),
    ("TT_sub", "T", "T", 40, 17, 23, 2.0, -0.25, 6, 1, 1, 3, 2, 6, 32, 32, 3),
    ("NN_blk", "N", "N", 64, 48, 70, 1.0, 1.0, 9, 17, 1, 9, 33, 1, 32, 16, 0),
]


def shapes(ta, tb, m, n, k, ia, ja, ib, jb, ic, jc, pad):
    ar, ac = (m if ta == "N" else k) + ia + 1, (k if ta == "N" else m) + ja + 1
    br, bc = (k if tb == "N" else n) + ib + 1, (n if tb == "N" else k) + jb + 1
    cr, cc = m + ic + 1, n + jc + 1
    return (ar + pad, ac), (br + pad, bc), (cr + pad, cc), (ar, ac), (br, bc), (cr, cc)


def sha(a):
    return hashlib.sha256(np.asfortranarray(a).tobytes(order="F")).hexdigest()


def main():
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "pdgemm_scalapack")
        lib = os.path.join(CONDA, "lib")
        subprocess.run(["gcc", "-O1", "-I", os.path.join(CONDA, "include"), "-o", exe,
                        os.path.join(HERE, "pdgemm_scalapack.c"), "-L", lib, "-Wl,--start-group",
                        "-lmkl_scalapack_lp64", "-lmkl_intel_lp64", "-lmkl_sequential", "-lmkl_core",
                        "-lmkl_blacs_intelmpi_lp64", "-Wl,--end-group", "-lmpi", "-lpthread", "-lm", "-ldl",
                        f"-Wl,-rpath,{lib}"], check=True)
        manifest = {}
        for case in CASES:
            name, ta, tb, m, n, k, alpha, beta, ia, ja, ib, jb, ic, jc, mb, nb, pad = case
            la, lb, lc, ga, gb, gc = shapes(ta, tb, m, n, k, ia, ja, ib, jb, ic, jc, pad)
            out = os.path.join(d, name + ".bin")
            subprocess.run([exe, ta, tb, *map(str, (m, n, k, alpha, beta, ia, ja, ib, jb, ic, jc, mb, nb, pad)),
                            out], check=True, timeout=60)
            C = np.fromfile(out, dtype=np.float64).reshape((lc[1], lc[0])).T
            A, B, C0 = oracle.uniform(*la, 1), oracle.uniform(*lb, 2), oracle.uniform(*lc, 3)
            path = os.path.join(HERE, f"pdgemm_{name}.npz")
            np.savez(path, case=np.array(json.dumps(case)), C=np.asfortranarray(C),
                     global_a=np.array(ga), global_b=np.array(gb), global_c=np.array(gc),
                     sha_A=sha(A), sha_B=sha(B), sha_C0=sha(C0))
            manifest[os.path.basename(path)] = {"case": case,
                                                "sha256": hashlib.sha256(open(path, "rb").read()).hexdigest()}
            print("wrote", path)
        with open(os.path.join(HERE, "pdgemm_manifest.json"), "w") as f:
            json.dump({"scalapack": "MKL ScaLAPACK lp64 + MPICH (/opt/conda)", "files": manifest}, f, indent=1)


if __name__ == "__main__":
    main()
