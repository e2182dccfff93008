"""Writes the golden fixtures under tests/golden/ (run once, in the build container).

The reference holds no golden vectors (its tests draw std::random_device
inputs and compare at run time).  These fixtures pin the single-rank path:

  gemm_<case>.npz   seeds of the inputs (SplitMix64, column-major) with a
                    SHA-256 of each regenerated input (inputs themselves are
                    stored only for the smallest cases), the host-BLAS result
                    (cblas_dgemm: the reference CPU path's arithmetic,
                    src/cosma/blas.cpp:35) and a double-double result.
  gemm_int_<case>.npz  seeds of integer-valued inputs whose products are
                    exact in FP64 (see int_matrix), with the exact result.
  transform_<case>.npz  block transform inputs and expected outputs
                    (alpha = 1, beta = 0 cases are bit-exact permutations).
  manifest.json     shapes, seeds, BLAS vendor and a SHA-256 of every file.

Shapes follow the reference's single-rank test cases: multiply_state(4,4,4,1)
and (3,4,5,1) (tests/multiply.cpp:144-145); pdgemm on a 1x1 grid 128x128x1280
TN, 83x83x83 NT, 83x77x83 TN (tests/pdgemm.cpp:420-456, :796-905); plus
100^3 and 512x32x736.
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
import oracle  # noqa: E402

GEMM_CASES = [
    # name, m, n, k, transa, transb, [(alpha, beta), ...]
    ("4x4x4_NN", 4, 4, 4, "N", "N", [(1.0, 0.0), (1.0, 1.0), (-1.3, 0.5)]),
    ("3x4x5_NN", 3, 4, 5, "N", "N", [(1.0, 0.0), (1.0, 1.0), (-1.3, 0.5)]),
    ("83x83x83_NT", 83, 83, 83, "N", "T", [(1.0, 0.0), (-1.3, 0.5), (1.0, 1.0)]),
    ("83x77x83_TN", 83, 77, 83, "T", "N", [(1.0, 0.0), (-1.3, 0.5), (1.0, 1.0)]),
    ("128x128x1280_TN", 128, 128, 1280, "T", "N", [(1.0, 0.0), (-1.3, 0.5)]),
    ("100x100x100_NN", 100, 100, 100, "N", "N", [(1.0, 0.0), (-1.3, 0.5), (1.0, 1.0)]),
    ("512x32x736_NN", 512, 32, 736, "N", "N", [(1.0, 0.0), (-1.3, 0.5)]),
    ("512x32x736_TT", 512, 32, 736, "T", "T", [(1.0, 1.0)]),
]

INT_CASES = [
    # integer entries in [-2^p, 2^p] with k * 2^(2p) < 2^53: every product and
    # partial sum is exact in FP64, so the result is known exactly.
    ("int_200x150x300_NN", 200, 150, 300, "N", "N", 16),
    ("int_129x257x1000_TN", 129, 257, 1000, "T", "N", 12),
]

TRANSFORM_CASES = [
    # name, op, rows, cols, lds_pad, ldd_pad, alpha, beta
    ("T_130x67", "T", 130, 67, 3, 0, 1.0, 0.0),
    ("N_130x67", "N", 130, 67, 5, 1, 1.0, 0.0),
    ("T_97x65_ab", "T", 97, 65, 0, 2, -1.3, 0.5),
    ("N_64x100_ab", "N", 64, 100, 0, 0, 2.0, 1.0),
    ("T_1x300", "T", 1, 300, 0, 0, 1.0, 0.0),
]


def stored_shape(t, rows, cols):
    """Column-major storage shape of X with op(X) rows x cols."""
    return (rows, cols) if t == "N" else (cols, rows)


def int_matrix(rows, cols, seed, p):
    """Column-major integer-valued matrix, entries in [-2^p, 2^p] from SplitMix64."""
    u = oracle.uniform(rows, cols, seed)           # uniform [-1, 1), exact dyadics
    return np.asfortranarray(np.round(u * 2 ** p))  # integers in [-2^p, 2^p]


def array_sha(a):
    return hashlib.sha256(np.asfortranarray(a).tobytes(order="F")).hexdigest()


def sha(path):
    h = hashlib.sha256()
    with open(path, "rb") as f:
        h.update(f.read())
    return h.hexdigest()


def main():
    oracle.hostblas.set_threads(1)
    manifest = {"blas": oracle.hostblas.vendor(), "rng": "splitmix64 uniform[-1,1)", "files": {}}
    seed = 1000
    for name, m, n, k, ta, tb, ab in GEMM_CASES:
        sa, sb, sc = seed, seed + 1, seed + 2
        seed += 3
        A = oracle.uniform(*stored_shape(ta, m, k), sa)
        B = oracle.uniform(*stored_shape(tb, k, n), sb)
        C0 = oracle.uniform(m, n, sc)
        out = {"seeds": np.array([sa, sb, sc]), "sha_A": array_sha(A), "sha_B": array_sha(B),
               "sha_C0": array_sha(C0), "m": m, "n": n, "k": k,
               "transa": ta, "transb": tb, "alphabeta": np.array(ab)}
        if m * n * k <= 64:
            out.update(A=A, B=B, C0=C0)
        for idx, (alpha, beta) in enumerate(ab):
            C = np.asfortranarray(C0.copy())
            oracle.hostblas.dgemm(ta, tb, m, n, k, alpha, A, A.shape[0], B, B.shape[0], beta, C, m)
            D = np.asfortranarray(C0.copy())
            oracle.dgemm(ta, tb, m, n, k, alpha, A, A.shape[0], B, B.shape[0], beta, D, m, dd=True)
            out[f"C_blas_{idx}"] = C
            if idx == 0:
                out[f"C_dd_{idx}"] = D
        path = os.path.join(HERE, f"gemm_{name}.npz")
        np.savez_compressed(path, **out)
        manifest["files"][os.path.basename(path)] = {"seeds": [sa, sb, sc], "sha256": sha(path)}

    for name, m, n, k, ta, tb, p in INT_CASES:
        sa, sb = seed, seed + 1
        seed += 2
        A = int_matrix(*stored_shape(ta, m, k), sa, p)
        B = int_matrix(*stored_shape(tb, k, n), sb, p)
        assert k * 2 ** (2 * p) < 2 ** 53
        Ai = A.astype(np.int64) if ta == "N" else A.T.astype(np.int64)
        Bi = B.astype(np.int64) if tb == "N" else B.T.astype(np.int64)
        exact = np.asfortranarray((Ai @ Bi).astype(np.float64))   # int64 arithmetic, exact
        path = os.path.join(HERE, f"gemm_{name}.npz")
        np.savez_compressed(path, seeds=np.array([sa, sb]), p=p, sha_A=array_sha(A), sha_B=array_sha(B),
                            C_exact=exact.astype(np.int64), m=m, n=n, k=k, transa=ta, transb=tb)
        manifest["files"][os.path.basename(path)] = {"sha256": sha(path)}

    for name, op, rows, cols, pad_s, pad_d, alpha, beta in TRANSFORM_CASES:
        sr, sc = (rows, cols) if op == "N" else (cols, rows)
        lds, ldd = sr + pad_s, rows + pad_d
        src = oracle.uniform(lds, sc, seed)
        dst0 = oracle.uniform(ldd, cols, seed + 1)
        seeds = np.array([seed, seed + 1])
        seed += 2
        dst = np.asfortranarray(dst0.copy())
        oracle.block_transform(op, rows, cols, alpha, src, lds, beta, dst, ldd)
        path = os.path.join(HERE, f"transform_{name}.npz")
        np.savez_compressed(path, seeds=seeds, sha_src=array_sha(src), sha_dst0=array_sha(dst0),
                            lds=lds, ldd=ldd, dst=dst, op=op, rows=rows, cols=cols,
                            alpha=alpha, beta=beta)
        manifest["files"][os.path.basename(path)] = {"sha256": sha(path)}

    with open(os.path.join(HERE, "manifest.json"), "w") as f:
        json.dump(manifest, f, indent=1, sort_keys=True)
    print("wrote", len(manifest["files"]), "fixtures")


if __name__ == "__main__":
    main()
