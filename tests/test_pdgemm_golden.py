"""ScaLAPACK's own pdgemm (1 x 1 grid, fixtures from make_pdgemm_golden.py)
against the oracle (CPU) and against cosma_pdgemm on the GPU.

The fixtures hold ScaLAPACK's whole output array: elements outside the C
submatrix must be untouched (bit for bit), and the submatrix must agree with
alpha * op(A_sub) op(B_sub) + beta * C_sub to the FP64 GEMM bound.  This pins
the meaning of ia/ja/ib/jb/ic/jc, the transposes and the descriptors
(src/cosma/cosma_pxgemm.cpp:15-388 follows ScaLAPACK's definition).
"""
import ctypes
import glob
import hashlib
import json
import os

import numpy as np
import pytest

import oracle
from helpers import GOLDEN, gemm_bound

FIXTURES = sorted(os.path.basename(p)[7:-4] for p in glob.glob(os.path.join(GOLDEN, "pdgemm_*.npz")))


def sha(a):
    return hashlib.sha256(np.asfortranarray(a).tobytes(order="F")).hexdigest()


def load(name):
    z = np.load(os.path.join(GOLDEN, f"pdgemm_{name}.npz"), allow_pickle=False)
    case = json.loads(str(z["case"]))
    _, ta, tb, m, n, k, alpha, beta, ia, ja, ib, jb, ic, jc, mb, nb, pad = case
    ga, gb, gc = (tuple(int(x) for x in z[key]) for key in ("global_a", "global_b", "global_c"))
    A = oracle.uniform(ga[0] + pad, ga[1], 1)
    B = oracle.uniform(gb[0] + pad, gb[1], 2)
    C0 = oracle.uniform(gc[0] + pad, gc[1], 3)
    assert sha(A) == str(z["sha_A"]) and sha(B) == str(z["sha_B"]) and sha(C0) == str(z["sha_C0"])
    return dict(ta=ta, tb=tb, m=m, n=n, k=k, alpha=alpha, beta=beta, ia=ia, ja=ja, ib=ib, jb=jb, ic=ic, jc=jc,
                mb=mb, nb=nb, A=A, B=B, C0=C0, ga=ga, gb=gb, gc=gc, C=np.asfortranarray(z["C"]))


def submatrix_reference(d):
    m, n, k = d["m"], d["n"], d["k"]
    ar, ac = (m, k) if d["ta"] == "N" else (k, m)
    br, bc = (k, n) if d["tb"] == "N" else (n, k)
    As = np.asfortranarray(d["A"][d["ia"] - 1:d["ia"] - 1 + ar, d["ja"] - 1:d["ja"] - 1 + ac])
    Bs = np.asfortranarray(d["B"][d["ib"] - 1:d["ib"] - 1 + br, d["jb"] - 1:d["jb"] - 1 + bc])
    ref = np.asfortranarray(d["C0"][d["ic"] - 1:d["ic"] - 1 + m, d["jc"] - 1:d["jc"] - 1 + n].copy())
    oracle.dgemm(d["ta"], d["tb"], m, n, k, d["alpha"], As, ar, Bs, br, d["beta"], ref, m, dd=True)
    return ref


def check(d, got):
    m, n = d["m"], d["n"]
    sub = (slice(d["ic"] - 1, d["ic"] - 1 + m), slice(d["jc"] - 1, d["jc"] - 1 + n))
    assert oracle.rel_fro(got[sub], submatrix_reference(d)) <= gemm_bound(d["k"])
    mask = np.ones(got.shape, dtype=bool)
    mask[sub] = False
    assert np.array_equal(got[mask], d["C0"][mask]), "elements outside the C submatrix changed"


def test_fixtures_present():
    assert len(FIXTURES) >= 5


@pytest.mark.parametrize("name", FIXTURES)
def test_oracle_agrees_with_scalapack(name):
    d = load(name)
    check(d, d["C"])


@pytest.mark.gpu
@pytest.mark.parametrize("name", FIXTURES)
def test_cosma_pdgemm_agrees_with_scalapack(name):
    import cosma_amd as ca
    d = load(name)
    lib = ca.lib()
    C = np.asfortranarray(d["C0"].copy())
    i = lambda v: ctypes.byref(ctypes.c_int(v))
    desc = lambda g, ld: (ctypes.c_int * 9)(1, 0, g[0], g[1], d["mb"], d["nb"], 0, 0, ld)
    pd = ctypes.POINTER(ctypes.c_double)
    lib.cosma_pdgemm(ctypes.c_char_p(d["ta"].encode()), ctypes.c_char_p(d["tb"].encode()), i(d["m"]), i(d["n"]),
                     i(d["k"]), ctypes.byref(ctypes.c_double(d["alpha"])), d["A"].ctypes.data_as(pd), i(d["ia"]),
                     i(d["ja"]), desc(d["ga"], d["A"].shape[0]), d["B"].ctypes.data_as(pd), i(d["ib"]), i(d["jb"]),
                     desc(d["gb"], d["B"].shape[0]), ctypes.byref(ctypes.c_double(d["beta"])),
                     C.ctypes.data_as(pd), i(d["ic"]), i(d["jc"]), desc(d["gc"], C.shape[0]))
    check(d, C)
    sub = (slice(d["ic"] - 1, d["ic"] - 1 + d["m"]), slice(d["jc"] - 1, d["jc"] - 1 + d["n"]))
    assert oracle.rel_fro(C[sub], d["C"][sub]) <= 2 * gemm_bound(d["k"])
