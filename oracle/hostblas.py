"""Host BLAS binding — the reference's CPU path on one rank (TEST INFRASTRUCTURE).

On one rank cosma::multiply makes exactly one local_multiply call
(src/cosma/multiply.cpp:369-394), which on a CPU build is
cosma::gemm -> cblas_dgemm(CblasColMajor, CblasNoTrans, CblasNoTrans, m, n, k,
alpha, A, m, B, k, beta, C, m) (src/cosma/local_multiply.cpp:369-371,
src/cosma/blas.cpp:24-48).  This module binds the host's cblas_dgemm so the
same call can be timed (the CPU baseline) and used as a large-size checker.

Probe order: MKL's single dynamic library (libmkl_rt), OpenBLAS, BLIS,
FlexiBLAS; else numpy's own BLAS through numpy.dot (vendor recorded).
Nothing here is loaded by the product library.
"""
import ctypes
import ctypes.util
import glob
import os

import numpy as np

CblasColMajor = 102
CblasNoTrans = 111
CblasTrans = 112

_CANDIDATES = [
    ("mkl", ["/opt/conda/lib/libmkl_rt.so.1", "/opt/conda/lib/libmkl_rt.so", "libmkl_rt.so.1", "libmkl_rt.so"]),
    ("openblas", ["libopenblas.so.0", "libopenblas.so"]),
    ("blis", ["libblis.so.4", "libblis.so"]),
    ("flexiblas", ["libflexiblas.so.3", "libflexiblas.so"]),
]

_state = {}


def _load():
    if "fn" in _state:
        return _state
    # MKL's default Intel OpenMP layer clashes with the GNU OpenMP runtime that
    # torch loads into the same process (threaded dgemm then returns wrong
    # values); the GNU layer shares torch's runtime.
    os.environ.setdefault("MKL_THREADING_LAYER", "GNU")
    for vendor, names in _CANDIDATES:
        for name in names:
            try:
                L = ctypes.CDLL(name)
                fn = L.cblas_dgemm
            except (OSError, AttributeError):
                continue
            fn.restype = None
            fn.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                           ctypes.c_double, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
                           ctypes.c_double, ctypes.c_void_p, ctypes.c_int]
            _state.update(fn=fn, vendor=vendor, path=name, lib=L)
            return _state
    _state.update(fn=None, vendor="numpy", path=np.__file__, lib=None)
    return _state


def vendor():
    s = _load()
    return f"{s['vendor']} ({s['path']})"


def set_threads(n):
    """Sets the BLAS thread count (MKL through mkl_set_num_threads when bound)."""
    s = _load()
    os.environ["OMP_NUM_THREADS"] = str(n)
    os.environ["MKL_NUM_THREADS"] = str(n)
    os.environ["OPENBLAS_NUM_THREADS"] = str(n)
    if s["vendor"] == "mkl":
        try:
            f = s["lib"].MKL_Set_Num_Threads
            f.argtypes = [ctypes.c_int]
            f(n)
        except AttributeError:
            pass


def dgemm(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc):
    """cblas_dgemm(ColMajor, op(A), op(B), ...) on column-major float64 arrays."""
    s = _load()
    if s["fn"] is not None:
        s["fn"](CblasColMajor, CblasNoTrans if ta == "N" else CblasTrans,
                CblasNoTrans if tb == "N" else CblasTrans, m, n, k, alpha,
                a.ctypes.data, lda, b.ctypes.data, ldb, beta, c.ctypes.data, ldc)
        return
    A = a[:m, :k] if ta == "N" else a[:k, :m].T
    B = b[:k, :n] if tb == "N" else b[:n, :k].T
    prod = alpha * (A @ B)
    if beta == 0.0:
        c[:m, :n] = prod
    else:
        c[:m, :n] = prod + beta * c[:m, :n]
