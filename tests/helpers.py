"""Shared helpers for the tests: fixture loading and host<->device transfer.

Matrices are column-major everywhere.  On the host they are numpy arrays in
Fortran order; on the device a column-major r x c matrix with leading
dimension ld is a torch tensor of shape (c, ld) (row j of the tensor is
column j of the matrix).
"""
import glob
import hashlib
import os

import numpy as np

import oracle

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
U = 2.0 ** -53


def array_sha(a):
    return hashlib.sha256(np.asfortranarray(a).tobytes(order="F")).hexdigest()


def stored_shape(t, rows, cols):
    return (rows, cols) if t == "N" else (cols, rows)


def int_matrix(rows, cols, seed, p):
    u = oracle.uniform(rows, cols, seed)
    return np.asfortranarray(np.round(u * 2 ** p))


def gemm_fixtures():
    return sorted(os.path.basename(p)[5:-4] for p in glob.glob(os.path.join(GOLDEN, "gemm_*.npz"))
                  if not os.path.basename(p).startswith("gemm_int_"))


def int_fixtures():
    return sorted(os.path.basename(p)[5:-4] for p in glob.glob(os.path.join(GOLDEN, "gemm_int_*.npz")))


def transform_fixtures():
    return sorted(os.path.basename(p)[10:-4] for p in glob.glob(os.path.join(GOLDEN, "transform_*.npz")))


def load_gemm(name):
    """Loads gemm_<name>.npz; regenerates and SHA-checks the inputs."""
    z = np.load(os.path.join(GOLDEN, f"gemm_{name}.npz"), allow_pickle=False)
    d = {k: z[k] for k in z.files}
    m, n, k = int(d["m"]), int(d["n"]), int(d["k"])
    ta, tb = str(d["transa"]), str(d["transb"])
    d.update(m=m, n=n, k=k, transa=ta, transb=tb)
    if name.startswith("int_"):
        sa, sb = (int(s) for s in d["seeds"])
        p = int(d["p"])
        d["A"] = int_matrix(*stored_shape(ta, m, k), sa, p)
        d["B"] = int_matrix(*stored_shape(tb, k, n), sb, p)
        d["C_exact"] = np.asfortranarray(d["C_exact"].astype(np.float64))
    else:
        sa, sb, sc = (int(s) for s in d["seeds"])
        A = oracle.uniform(*stored_shape(ta, m, k), sa)
        B = oracle.uniform(*stored_shape(tb, k, n), sb)
        C0 = oracle.uniform(m, n, sc)
        if "A" in d:
            assert np.array_equal(d["A"], A) and np.array_equal(d["B"], B) and np.array_equal(d["C0"], C0)
        d.update(A=A, B=B, C0=C0)
        d["alphabeta"] = [tuple(x) for x in d["alphabeta"]]
    assert array_sha(d["A"]) == str(d["sha_A"]), "regenerated A differs from the fixture"
    assert array_sha(d["B"]) == str(d["sha_B"]), "regenerated B differs from the fixture"
    if "sha_C0" in d:
        assert array_sha(d["C0"]) == str(d["sha_C0"])
    return d


def load_transform(name):
    z = np.load(os.path.join(GOLDEN, f"transform_{name}.npz"), allow_pickle=False)
    d = {k: z[k] for k in z.files}
    op, rows, cols = str(d["op"]), int(d["rows"]), int(d["cols"])
    lds, ldd = int(d["lds"]), int(d["ldd"])
    s0, s1 = (int(s) for s in d["seeds"])
    sc = cols if op == "N" else rows
    src = oracle.uniform(lds, sc, s0)
    dst0 = oracle.uniform(ldd, cols, s1)
    assert array_sha(src) == str(d["sha_src"]) and array_sha(dst0) == str(d["sha_dst0"])
    d.update(op=op, rows=rows, cols=cols, lds=lds, ldd=ldd, src=src, dst0=dst0,
             alpha=float(d["alpha"]), beta=float(d["beta"]))
    return d


def to_dev(a, ld=None):
    """Column-major numpy matrix -> device tensor of shape (cols, ld); returns (tensor, ld)."""
    import torch
    a = np.asfortranarray(a)
    rows, cols = a.shape
    ld = rows if ld is None else ld
    t = torch.zeros((cols, ld), dtype=torch.float64)
    t[:, :rows] = torch.from_numpy(np.ascontiguousarray(a.T))
    return t.cuda(), ld


def from_dev(t, rows):
    """Device tensor (cols, ld) -> column-major numpy rows x cols."""
    return np.asfortranarray(t[:, :rows].cpu().numpy().T)


def gemm_bound(k):
    """Acceptance bound on the relative Frobenius error of an FP64 GEMM."""
    return 4.0 * max(k, 1) * U
