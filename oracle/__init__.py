"""CPU oracle for the single-rank local-multiply path — TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg import
this package, and only as the checker.  The product (cosma_amd) never imports
it and has no CPU fallback.

Contents
  * liboracle.so (oracle.c): restatements of the reference's CPU arithmetic,
    each citing the reference file:line it follows.
  * _ref/liblm_ref.so: the reference's own local_multiply.cpp compiled
    unmodified (when it was built); used to pin the naive-loop restatement.
  * hostblas: the reference's CPU path on one rank is a single host-BLAS
    cblas_dgemm call (src/cosma/blas.cpp:35); this binds the host's BLAS
    (MKL's libmkl_rt if present) for large checks and the CPU baseline.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "liblm_ref.so")

_i64 = ctypes.c_int64
_dbl = ctypes.c_double
_vp = ctypes.c_void_p
_lib = None
_ref = None


def build():
    """Compiles liboracle.so (and the reference _ref when its tree is present)."""
    import subprocess
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        L.oracle_splitmix64.restype = ctypes.c_uint64
        L.oracle_splitmix64.argtypes = [ctypes.c_uint64]
        L.oracle_fill_uniform.restype = None
        L.oracle_fill_uniform.argtypes = [_vp, _i64, ctypes.c_uint64, ctypes.c_uint64]
        L.oracle_local_multiply_cpu.restype = None
        L.oracle_local_multiply_cpu.argtypes = [_vp, _vp, _vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dbl, _dbl]
        for name in ("oracle_dgemm", "oracle_dgemm_dd"):
            f = getattr(L, name)
            f.restype = ctypes.c_int
            f.argtypes = [ctypes.c_char, ctypes.c_char, _i64, _i64, _i64, _dbl, _vp, _i64, _vp, _i64, _dbl, _vp, _i64]
        L.oracle_block_transform.restype = ctypes.c_int
        L.oracle_block_transform.argtypes = [ctypes.c_char, _i64, _i64, _dbl, _vp, _i64, _dbl, _vp, _i64]
        L.oracle_numroc.restype = _i64
        L.oracle_numroc.argtypes = [_i64, _i64, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.oracle_global_to_local.restype = None
        L.oracle_global_to_local.argtypes = [_i64, _i64, ctypes.c_int, ctypes.c_int,
                                             ctypes.POINTER(ctypes.c_int), ctypes.POINTER(_i64)]
        L.oracle_block_cyclic_gather.restype = None
        L.oracle_block_cyclic_gather.argtypes = [_vp, _i64, _i64, _i64, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                                 ctypes.c_int, ctypes.c_int, ctypes.c_int, _i64, _i64, _i64, _i64,
                                                 _vp, _i64]
        _lib = L
    return _lib


def ref_lib():
    """The reference's compiled local_multiply.cpp, or None if not built."""
    global _ref
    if _ref is None and os.path.exists(REF_LIB):
        L = ctypes.CDLL(REF_LIB, mode=os.RTLD_LAZY | os.RTLD_LOCAL)
        f = L._ZN5cosma18local_multiply_cpuIdEEvPT_S2_S2_iiiS1_S1_
        f.restype = None
        f.argtypes = [_vp, _vp, _vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dbl, _dbl]
        _ref = f
    return _ref


def _p(a):
    assert a.flags["F_CONTIGUOUS"] or a.ndim == 1
    return a.ctypes.data


# ------------------------------------------------------------------ helpers
def splitmix64(z):
    return lib().oracle_splitmix64(z)


def uniform(rows, cols, seed, offset=0):
    """Column-major rows x cols matrix of uniform [-1, 1) values (SplitMix64)."""
    x = np.empty(rows * cols, dtype=np.float64)
    lib().oracle_fill_uniform(x.ctypes.data, x.size, seed, offset)
    return x.reshape((rows, cols), order="F")


def uniform_at(seed, counters):
    """Uniform [-1, 1) values at arbitrary stream positions (numpy, vectorised).

    Same formula as oracle_fill_uniform: u(splitmix64(seed * golden + counter)).
    """
    with np.errstate(over="ignore"):
        z = np.asarray(counters, dtype=np.uint64) + np.uint64((seed * 0x9E3779B97F4A7C15) % 2 ** 64)
        z = z + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * 2.0 ** -53 * 2.0 - 1.0


def local_multiply_cpu(a, b, c, m, n, k, alpha, beta):
    """Naive loop of local_multiply.cpp:277-297 on column-major arrays (in place on c)."""
    lib().oracle_local_multiply_cpu(_p(a), _p(b), _p(c), m, n, k, alpha, beta)


def local_multiply_cpu_ref(a, b, c, m, n, k, alpha, beta):
    """Same, through the reference's own compiled code (None if unavailable)."""
    f = ref_lib()
    if f is None:
        return None
    f(_p(a), _p(b), _p(c), m, n, k, alpha, beta)
    return True


def dgemm(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, dd=False):
    """BLAS-semantics dgemm on column-major numpy arrays (in place on c)."""
    f = lib().oracle_dgemm_dd if dd else lib().oracle_dgemm
    rc = f(ta.encode(), tb.encode(), m, n, k, alpha, _p(a), lda, _p(b), ldb, beta, _p(c), ldc)
    if rc:
        raise ValueError("oracle_dgemm: invalid arguments")


def dgemm_raw(ta, tb, m, n, k, alpha, pa, lda, pb, ldb, beta, pc, ldc, dd=False):
    """As dgemm, on raw host addresses (for sub-blocks of large host buffers)."""
    f = lib().oracle_dgemm_dd if dd else lib().oracle_dgemm
    rc = f(ta.encode(), tb.encode(), m, n, k, alpha, pa, lda, pb, ldb, beta, pc, ldc)
    if rc:
        raise ValueError("oracle_dgemm: invalid arguments")


def block_transform(op, rows, cols, alpha, src, lds, beta, dst, ldd):
    rc = lib().oracle_block_transform(op.encode(), rows, cols, alpha, _p(src), lds, beta, _p(dst), ldd)
    if rc:
        raise ValueError("oracle_block_transform: invalid arguments")


def numroc(n, nb, proc, src, nprocs):
    return lib().oracle_numroc(n, nb, proc, src, nprocs)


def global_to_local(g, nb, src, nprocs):
    p = ctypes.c_int()
    l = _i64()
    lib().oracle_global_to_local(g, nb, src, nprocs, ctypes.byref(p), ctypes.byref(l))
    return p.value, l.value


def block_cyclic_gather(local, lld, mb, nb, rsrc, csrc, nprow, npcol, myrow, mycol, ia, ja, m, n, dst, ldd):
    lib().oracle_block_cyclic_gather(_p(local), lld, mb, nb, rsrc, csrc, nprow, npcol, myrow, mycol,
                                     ia, ja, m, n, _p(dst), ldd)


def rel_fro(got, ref):
    """Relative Frobenius error ||got - ref||_F / ||ref||_F (0 if both zero)."""
    d = np.linalg.norm((np.asarray(got) - np.asarray(ref)).ravel())
    r = np.linalg.norm(np.asarray(ref).ravel())
    return float(d / r) if r > 0 else float(d)


from . import hostblas  # noqa: E402,F401
