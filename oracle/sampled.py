"""Checks of full-size results against the double-double oracle on sampled
rows and columns of C — TEST INFRASTRUCTURE ONLY (tests/ and bench.py's
error checks).

At the benchmark sizes (16384^3, k = 2^20) the double-double oracle cannot
compute all of C, so a full-size check takes whole rows and whole columns of
op(A) op(B) from the oracle (each entry with all k terms), and compares the
same lines of the GPU result; the rest of C is checked for finiteness and
against the plain FP64 GPU DGEMM (itself parity-tested against the oracle at
small sizes).  The oracle runs on several host threads (ctypes releases the
interpreter lock during the calls).
"""
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from . import dgemm_raw


def host_threads():
    return max(1, min(16, len(os.sched_getaffinity(0))))


def _a_row(ta, pa, lda, i):
    """Address of op(A)(i, 0) for a 1-row slice with the same lda."""
    return pa + 8 * (i if ta == "N" else i * lda)


def _b_col(tb, pb, ldb, j):
    return pb + 8 * (j * ldb if tb == "N" else j)


def dd_lines(ta, tb, m, n, k, pa, lda, pb, ldb, rows, cols, threads=None):
    """Double-double op(A) op(B) at whole rows `rows` (each length n) and
    whole columns `cols` (each length m); A and B are host addresses."""
    threads = threads or host_threads()
    row_out = {i: np.zeros(n) for i in rows}
    col_out = {j: np.zeros(m) for j in cols}
    tasks = []
    nc = max(1, -(-n // threads))
    for i in rows:
        for j0 in range(0, n, nc):
            c = min(nc, n - j0)
            tasks.append((ta, tb, 1, c, k, 1.0, _a_row(ta, pa, lda, i), lda, _b_col(tb, pb, ldb, j0), ldb, 0.0,
                          row_out[i].ctypes.data + 8 * j0, 1))
    mc = max(1, -(-m // threads))
    for j in cols:
        for i0 in range(0, m, mc):
            r = min(mc, m - i0)
            tasks.append((ta, tb, r, 1, k, 1.0, _a_row(ta, pa, lda, i0), lda, _b_col(tb, pb, ldb, j), ldb, 0.0,
                          col_out[j].ctypes.data + 8 * i0, r))
    with ThreadPoolExecutor(threads) as ex:
        list(ex.map(lambda t: dgemm_raw(*t, dd=True), tasks))
    return row_out, col_out


def lines_error(got_rows, got_cols, ref_rows, ref_cols):
    """Relative Frobenius error over the sampled lines."""
    num = den = 0.0
    for key, got in got_rows.items():
        num += float(((got - ref_rows[key]) ** 2).sum())
        den += float((ref_rows[key] ** 2).sum())
    for key, got in got_cols.items():
        num += float(((got - ref_cols[key]) ** 2).sum())
        den += float((ref_cols[key] ** 2).sum())
    return (num / den) ** 0.5 if den > 0 else num ** 0.5
