"""The C++ COSMA interface on the GPU: local_multiply (three overloads and
local_multiply_cpu, the reference's test-utility name),
multiply with CosmaMatrix and sequential steps, the ScaLAPACK entry points
(prefixed and interposed) and the C layout API, each against the
double-double oracle.

The C++ driver (tests/cpp/front_test) generates its inputs with the same
counter-based generator as oracle.uniform, so the oracle sees identical
values.  Untouched parts of the ScaLAPACK arrays must come back bit for bit.
"""
import ctypes
import os
import subprocess
import tempfile

import numpy as np
import pytest

import oracle
from helpers import gemm_bound
from cosma_amd import build as cbuild

pytestmark = pytest.mark.gpu


def drive(*args):
    if not os.path.exists(cbuild.TEST_DRIVER):
        cbuild.build_test_driver()
    with tempfile.TemporaryDirectory() as d:
        out = os.path.join(d, "out.bin")
        print("front_test", *args, flush=True)
        r = subprocess.run([cbuild.TEST_DRIVER, *map(str, args), out], capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        return np.fromfile(out, dtype=np.float64)


def expected(ta, tb, m, n, k, alpha, A, B, beta, C):
    ref = np.asfortranarray(C.copy())
    oracle.dgemm(ta, tb, m, n, k, alpha, A, A.shape[0], B, B.shape[0], beta, ref, m, dd=True)
    return ref


@pytest.mark.parametrize("variant", ["singleton", "context", "handle", "cpu_name"])
@pytest.mark.parametrize("m,n,k,alpha,beta", [(200, 150, 170, 1.0, 0.0), (333, 129, 257, -1.3, 0.5)])
def test_local_multiply(variant, m, n, k, alpha, beta):
    got = drive("local_multiply", variant, m, n, k, alpha, beta).reshape((n, m)).T
    A, B, C = oracle.uniform(m, k, 1), oracle.uniform(k, n, 2), oracle.uniform(m, n, 3)
    assert oracle.rel_fro(got, expected("N", "N", m, n, k, alpha, A, B, beta, C)) <= gemm_bound(k)


@pytest.mark.parametrize("steps", ["-", "sm2", "sk3", "sm2,sk2", "sn3,sm2", "sk2,sn2,sm3"])
def test_multiply_with_sequential_steps(steps):
    """k-splits accumulate with beta = 1 after the first chunk (src/cosma/multiply.cpp:550-569)."""
    m, n, k = 97, 83, 131
    got = drive("multiply", m, n, k, steps, -0.75, 0.5).reshape((n, m)).T
    A, B, C = oracle.uniform(m, k, 1), oracle.uniform(k, n, 2), oracle.uniform(m, n, 3)
    assert oracle.rel_fro(got, expected("N", "N", m, n, k, -0.75, A, B, 0.5, C)) <= gemm_bound(k)


@pytest.mark.parametrize("ta,tb", [("N", "N"), ("T", "N"), ("N", "T"), ("T", "T")])
@pytest.mark.parametrize("sub", [(1, 1, 1, 1, 1, 1), (3, 2, 2, 4, 5, 3)])
def test_cosma_pdgemm_submatrices(ta, tb, sub):
    m, n, k, alpha, beta = 61, 45, 77, 1.25, -0.5
    ia, ja, ib, jb, ic, jc = sub
    got = drive("pdgemm", ta, tb, m, n, k, alpha, beta, ia, ja, ib, jb, ic, jc)
    na, nb = ta == "N", tb == "N"
    ar, ac = (m if na else k) + ia + 1, (k if na else m) + ja + 1
    br, bc = (k if nb else n) + ib + 1, (n if nb else k) + jb + 1
    cr, cc = m + ic + 1, n + jc + 1
    lda, ldb, ldc = ar + 3, br + 3, cr + 3
    Af, Bf, Cf = oracle.uniform(lda, ac, 1), oracle.uniform(ldb, bc, 2), oracle.uniform(ldc, cc, 3)
    got = got.reshape((cc, ldc)).T
    Asub = Af[ia - 1:ia - 1 + (m if na else k), ja - 1:ja - 1 + (k if na else m)]
    Bsub = Bf[ib - 1:ib - 1 + (k if nb else n), jb - 1:jb - 1 + (n if nb else k)]
    Csub = Cf[ic - 1:ic - 1 + m, jc - 1:jc - 1 + n]
    ref = expected(ta, tb, m, n, k, alpha, np.asfortranarray(Asub), np.asfortranarray(Bsub), beta, Csub)
    out_sub = got[ic - 1:ic - 1 + m, jc - 1:jc - 1 + n]
    assert oracle.rel_fro(out_sub, ref) <= gemm_bound(k)
    mask = np.ones_like(Cf, dtype=bool)
    mask[ic - 1:ic - 1 + m, jc - 1:jc - 1 + n] = False
    assert np.array_equal(got[mask], Cf[mask]), "elements outside the C submatrix changed"


# ------------------------------------------------------------ ctypes callers

PI = ctypes.POINTER(ctypes.c_int)
PD = ctypes.POINTER(ctypes.c_double)


def call_pdgemm(fn, ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, ia=1, ja=1, ib=1, jb=1, ic=1, jc=1):
    desc = lambda rows, cols, ld: (ctypes.c_int * 9)(1, 0, rows, cols, 64, 64, 0, 0, ld)
    i = lambda v: ctypes.byref(ctypes.c_int(v))
    fn(ctypes.c_char_p(ta.encode()), ctypes.c_char_p(tb.encode()), i(m), i(n), i(k),
       ctypes.byref(ctypes.c_double(alpha)), A.ctypes.data_as(PD), i(ia), i(ja), desc(A.shape[0], A.shape[1], lda),
       B.ctypes.data_as(PD), i(ib), i(jb), desc(B.shape[0], B.shape[1], ldb), ctypes.byref(ctypes.c_double(beta)),
       C.ctypes.data_as(PD), i(ic), i(jc), desc(C.shape[0], C.shape[1], ldc))


@pytest.mark.parametrize("symbol", ["pdgemm", "pdgemm_", "PDGEMM", "PDGEMM_"])
def test_interposed_scalapack_symbols(symbol):
    """libcosma_amd_pxgemm.so's pdgemm spellings, called as a Fortran/C caller would."""
    import torch   # the HIP runtime torch bundles must load first
    torch.cuda.init()
    lib = ctypes.CDLL(cbuild.PXGEMM_LIB)
    m, n, k = 130, 70, 90
    A, B = oracle.uniform(k, m, 5), oracle.uniform(k, n, 6)        # op(A) = A^T
    C0 = oracle.uniform(m, n, 7)
    C = np.asfortranarray(C0.copy())
    call_pdgemm(getattr(lib, symbol), "T", "N", m, n, k, 2.0, A, k, B, k, 1.0, C, m)
    assert oracle.rel_fro(C, expected("T", "N", m, n, k, 2.0, A, B, 1.0, C0)) <= gemm_bound(k)


def test_pdgemm_corner_cases_through_prefixed_symbol():
    import cosma_amd as ca
    lib = ca.lib()
    m, n, k = 40, 30, 20
    A, B = oracle.uniform(m, k, 1), oracle.uniform(k, n, 2)
    # k == 0: C <- beta * C
    C0 = oracle.uniform(m, n, 3)
    C = np.asfortranarray(C0.copy())
    call_pdgemm(lib.cosma_pdgemm, "N", "N", m, n, 0, 1.0, A, m, B, k, 0.25, C, m)
    assert np.array_equal(C, 0.25 * C0)
    # beta == 0 never reads C
    C = np.full((m, n), np.nan, order="F")
    call_pdgemm(lib.cosma_pdgemm, "N", "N", m, n, k, 1.0, A, m, B, k, 0.0, C, m)
    assert np.isfinite(C).all()
    assert oracle.rel_fro(C, expected("N", "N", m, n, k, 1.0, A, B, 0.0, np.zeros((m, n)))) <= gemm_bound(k)
    # m == 0: nothing happens
    C = np.asfortranarray(C0.copy())
    call_pdgemm(lib.cosma_pdgemm, "N", "N", 0, n, k, 1.0, A, m, B, k, 0.0, C, m)
    assert np.array_equal(C, C0)


class Block(ctypes.Structure):
    _fields_ = [("data", ctypes.c_void_p), ("ld", ctypes.c_int), ("row", ctypes.c_int), ("col", ctypes.c_int)]


class Layout(ctypes.Structure):
    _fields_ = [("rowblocks", ctypes.c_int), ("colblocks", ctypes.c_int), ("rowsplit", PI), ("colsplit", PI),
                ("owners", PI), ("nlocalblocks", ctypes.c_int), ("localblocks", ctypes.POINTER(Block))]


def make_layout(X, rsplit, csplit, keep):
    """Grid layout of X with each block in its own Fortran array (ld = rows + 1)."""
    bufs, blocks = [], []
    for bi in range(len(rsplit) - 1):
        for bj in range(len(csplit) - 1):
            sub = X[rsplit[bi]:rsplit[bi + 1], csplit[bj]:csplit[bj + 1]]
            buf = np.zeros((sub.shape[0] + 1, max(sub.shape[1], 1)), order="F")
            buf[:sub.shape[0], :sub.shape[1]] = sub
            bufs.append(buf)
            blocks.append(Block(buf.ctypes.data, buf.shape[0], bi, bj))
    rs = (ctypes.c_int * len(rsplit))(*rsplit)
    cs = (ctypes.c_int * len(csplit))(*csplit)
    owners = (ctypes.c_int * (len(blocks)))(*([0] * len(blocks)))
    barr = (Block * len(blocks))(*blocks)
    lay = Layout(len(rsplit) - 1, len(csplit) - 1, rs, cs, owners, len(blocks), barr)
    keep += [bufs, rs, cs, owners, barr]
    return lay, bufs


def gather(bufs, rsplit, csplit, shape):
    out = np.zeros(shape)
    it = iter(bufs)
    for bi in range(len(rsplit) - 1):
        for bj in range(len(csplit) - 1):
            b = next(it)
            out[rsplit[bi]:rsplit[bi + 1], csplit[bj]:csplit[bj + 1]] = \
                b[:rsplit[bi + 1] - rsplit[bi], :csplit[bj + 1] - csplit[bj]]
    return out


@pytest.mark.parametrize("ta,tb", [("N", "N"), ("T", "N")])
def test_dmultiply_using_layout(ta, tb):
    import cosma_amd as ca
    lib = ca.lib()
    m, n, k = 90, 70, 110
    A = oracle.uniform(m, k, 1) if ta == "N" else oracle.uniform(k, m, 1)
    B = oracle.uniform(k, n, 2)
    C0 = oracle.uniform(m, n, 3)
    keep = []
    ra = [0, 30, 31, A.shape[0]]
    ca_ = [0, 50, A.shape[1]]
    la, _ = make_layout(A, ra, ca_, keep)
    lb, _ = make_layout(B, [0, 64, k], [0, 7, 40, n], keep)
    rc, cc = [0, 45, m], [0, 35, n]
    lc, cbufs = make_layout(C0, rc, cc, keep)
    lib.dmultiply_using_layout(ctypes.c_int(0), ctypes.c_char_p(ta.encode()), ctypes.c_char_p(tb.encode()),
                               ctypes.byref(ctypes.c_double(-1.5)), ctypes.byref(la), ctypes.byref(lb),
                               ctypes.byref(ctypes.c_double(0.5)), ctypes.byref(lc))
    got = gather(cbufs, rc, cc, (m, n))
    assert oracle.rel_fro(got, expected(ta, tb, m, n, k, -1.5, A, B, 0.5, C0)) <= gemm_bound(k)


def _driver_env(args, env_extra):
    env = {k: v for k, v in os.environ.items() if not k.startswith("COSMA_")}
    env.update(env_extra)
    print("front_test", *args, env_extra, flush=True)
    r = subprocess.run([cbuild.TEST_DRIVER, *args], capture_output=True, text=True, timeout=120, env=env)
    assert r.returncode == 0, (r.stdout, r.stderr)
    return r.stdout.splitlines()


def test_knobs_reach_the_device_handle():
    """COSMA_GPU_STREAMS and COSMA_GPU_MAX_TILE_{M,N,K} become the streams
    and streaming tiles of the context's device handle."""
    lines = _driver_env(["env", "handle"], {"COSMA_GPU_STREAMS": "3", "COSMA_GPU_MAX_TILE_M": "700",
                                            "COSMA_GPU_MAX_TILE_N": "900", "COSMA_GPU_MAX_TILE_K": "1100"})
    assert "handle 0 3 700 900 1100" in lines, lines
    assert "handle 0 2 5000 5000 5000" in _driver_env(["env", "handle"], {})


def test_memory_pool_page_locked():
    """With COSMA_GPU_MEMORY_PINNING on (the default) the pool's storage is
    page-locked host memory, with the same stack and growth behaviour."""
    lines = _driver_env(["pool"], {})
    assert not [l for l in lines if l.startswith("FAIL")], lines
    assert [l for l in lines if l.startswith("pool capacity")][0].endswith("page_locked 1")


def test_multiply_from_pageable_pool():
    """COSMA_GPU_MEMORY_PINNING=OFF: matrices in pageable memory, staged by the
    streaming GEMM; same results as from the page-locked pool."""
    import tempfile as tf
    m, n, k = 150, 130, 170
    with tf.TemporaryDirectory() as d:
        outs = []
        for pin in ("ON", "OFF"):
            out = os.path.join(d, f"c_{pin}.bin")
            _driver_env(["multiply", str(m), str(n), str(k), "sm2,sk2", "1.0", "0.5", out],
                        {"COSMA_GPU_MEMORY_PINNING": pin})
            outs.append(np.fromfile(out, dtype=np.float64))
    assert np.array_equal(outs[0], outs[1])
