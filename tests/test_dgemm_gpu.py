"""Parity of the FP64 MFMA DGEMM (through the C ABI) with the oracle.

Acceptance (stated by the task for FP64 results): relative Frobenius error
against the oracle <= 4 * k * 2^-53.  Integer-valued inputs whose products
and partial sums are exact in FP64 must come out bit-exact.  Semantics from
the reference: beta == 0 never reads C (utils/pxgemm_utils.hpp:604-616),
k == 0 or alpha == 0 gives beta * C, m == 0 or n == 0 is a no-op
(src/cosma/multiply.cpp:99-106)."""
import numpy as np
import pytest

import oracle
from helpers import (from_dev, gemm_bound, gemm_fixtures, int_fixtures, load_gemm, stored_shape, to_dev)

pytestmark = pytest.mark.gpu


def run_gpu(handle, ta, tb, m, n, k, alpha, A, B, beta, C0, lda=None, ldb=None, ldc=None, nan_c=False):
    import torch
    dA, lda = to_dev(A, lda)
    dB, ldb = to_dev(B, ldb)
    dC, ldc = to_dev(C0, ldc)
    if nan_c:
        dC.fill_(float("nan"))
    handle.dgemm(ta, tb, m, n, k, alpha, dA.data_ptr(), lda, dB.data_ptr(), ldb, beta, dC.data_ptr(), ldc)
    torch.cuda.synchronize()
    return from_dev(dC, m)


def dd_ref(ta, tb, m, n, k, alpha, A, B, beta, C0):
    C = C0.copy(order="F")
    oracle.dgemm(ta, tb, m, n, k, alpha, A, A.shape[0], B, B.shape[0], beta, C, m, dd=True)
    return C


@pytest.mark.parametrize("name", gemm_fixtures())
def test_golden_fixtures(handle, name):
    d = load_gemm(name)
    m, n, k = d["m"], d["n"], d["k"]
    for idx, (alpha, beta) in enumerate(d["alphabeta"]):
        got = run_gpu(handle, d["transa"], d["transb"], m, n, k, alpha, d["A"], d["B"], beta, d["C0"],
                      nan_c=(beta == 0.0))
        assert np.isfinite(got).all()
        err = oracle.rel_fro(got, d[f"C_blas_{idx}"])
        assert err <= gemm_bound(k), (name, alpha, beta, err)
        if f"C_dd_{idx}" in d:
            assert oracle.rel_fro(got, d[f"C_dd_{idx}"]) <= gemm_bound(k)


@pytest.mark.parametrize("name", int_fixtures())
def test_integer_golden_is_bit_exact(handle, name):
    d = load_gemm(name)
    m, n, k = d["m"], d["n"], d["k"]
    got = run_gpu(handle, d["transa"], d["transb"], m, n, k, 1.0, d["A"], d["B"], 0.0, np.zeros((m, n), order="F"),
                  nan_c=True)
    assert np.array_equal(got, d["C_exact"])


SHAPES = [(1, 1, 1), (3, 4, 5), (4, 4, 4), (17, 31, 83), (128, 128, 16), (129, 257, 100), (300, 200, 513),
          (258, 130, 17), (511, 383, 64), (1, 300, 7), (300, 1, 7), (64, 64, 1)]


@pytest.mark.parametrize("m,n,k", SHAPES)
@pytest.mark.parametrize("ta,tb", [("N", "N"), ("T", "N"), ("N", "T"), ("T", "T")])
def test_random_shapes_all_transposes(handle, m, n, k, ta, tb):
    A = oracle.uniform(*stored_shape(ta, m, k), m + 7 * n + k)
    B = oracle.uniform(*stored_shape(tb, k, n), 2 * m + n + 3 * k)
    C0 = oracle.uniform(m, n, 99)
    for alpha, beta in ((1.0, 0.0), (-1.3, 0.5)):
        got = run_gpu(handle, ta, tb, m, n, k, alpha, A, B, beta, C0, nan_c=(beta == 0.0))
        ref = dd_ref(ta, tb, m, n, k, alpha, A, B, beta, C0)
        assert oracle.rel_fro(got, ref) <= gemm_bound(k), (m, n, k, ta, tb, alpha, beta)


@pytest.mark.parametrize("cfg", [1, 2, 3])
def test_every_tile_configuration(handle, cfg):
    import cosma_amd as ca
    try:
        ca.set_dgemm_tuning(cfg, 0)
        for (m, n, k, ta, tb) in [(300, 270, 129, "T", "N"), (257, 513, 64, "N", "N"), (129, 65, 300, "N", "T")]:
            A = oracle.uniform(*stored_shape(ta, m, k), 1)
            B = oracle.uniform(*stored_shape(tb, k, n), 2)
            C0 = oracle.uniform(m, n, 3)
            got = run_gpu(handle, ta, tb, m, n, k, 0.75, A, B, 1.0, C0)
            assert oracle.rel_fro(got, dd_ref(ta, tb, m, n, k, 0.75, A, B, 1.0, C0)) <= gemm_bound(k)
    finally:
        ca.set_dgemm_tuning(0, 0)


@pytest.mark.parametrize("m,n,k", [(128, 128, 16), (256, 384, 64), (640, 512, 1040), (1024, 1152, 208)])
@pytest.mark.parametrize("ta,tb", [("T", "N"), ("N", "N"), ("N", "T"), ("T", "T")])
def test_direct_lds_kernel_is_bit_identical(handle, m, n, k, ta, tb):
    """Tile configuration 4 (direct global-to-LDS loads, 128 x 128 tiles, two
    workgroups per CU) sums every element in the same order as the other
    tile shapes: its results must equal theirs bit for bit, and the oracle's
    within the FP64 bound."""
    import cosma_amd as ca
    A = oracle.uniform(*stored_shape(ta, m, k), 11)
    B = oracle.uniform(*stored_shape(tb, k, n), 12)
    C0 = oracle.uniform(m, n, 13)
    results = {}
    try:
        for cfg in (4, 3, 1):
            ca.set_dgemm_tuning(cfg, 0)
            for alpha, beta in ((1.0, 0.5), (-0.75, 0.0)):
                results[(cfg, beta)] = run_gpu(handle, ta, tb, m, n, k, alpha, A, B, beta, C0, nan_c=(beta == 0.0))
    finally:
        ca.set_dgemm_tuning(0, 0)
    for beta, alpha in ((0.5, 1.0), (0.0, -0.75)):
        assert np.array_equal(results[(4, beta)], results[(3, beta)])
        assert np.array_equal(results[(4, beta)], results[(1, beta)])
        ref = dd_ref(ta, tb, m, n, k, alpha, A, B, beta, C0)
        assert oracle.rel_fro(results[(4, beta)], ref) <= gemm_bound(k)


@pytest.mark.parametrize("ta,tb", [("T", "N"), ("N", "N"), ("N", "T"), ("T", "T")])
@pytest.mark.parametrize("cpad", [4, 3])
def test_direct_lds_kernel_padded_leading_dimensions(handle, ta, tb, cpad):
    """Even padded leading dimensions of A and B keep the direct-LDS kernel
    (16-byte aligned rows and k-rows) and must give the same bits as the
    other tile shapes; an odd ldc takes its 8-byte epilogue."""
    import cosma_amd as ca
    m, n, k = 256, 384, 96
    A = oracle.uniform(*stored_shape(ta, m, k), 21)
    B = oracle.uniform(*stored_shape(tb, k, n), 22)
    C0 = oracle.uniform(m, n, 23)
    lda, ldb, ldc = A.shape[0] + 6, B.shape[0] + 2, m + cpad
    out = {}
    try:
        for cfg in (4, 1):
            ca.set_dgemm_tuning(cfg, 0)
            out[cfg] = run_gpu(handle, ta, tb, m, n, k, 1.25, A, B, -0.5, C0, lda=lda, ldb=ldb, ldc=ldc)
    finally:
        ca.set_dgemm_tuning(0, 0)
    assert np.array_equal(out[4], out[1])
    assert oracle.rel_fro(out[4], dd_ref(ta, tb, m, n, k, 1.25, A, B, -0.5, C0)) <= gemm_bound(k)


def test_direct_lds_kernel_falls_back_off_its_shapes(handle):
    """Configuration 4 on shapes outside its conditions (m not a multiple of
    128, k not of 16, an odd leading dimension) takes the automatic kernel
    and stays correct."""
    import cosma_amd as ca
    try:
        ca.set_dgemm_tuning(4, 0)
        for (m, n, k, ta, tb, pad) in [(300, 256, 64, "T", "N", 0), (256, 256, 40, "N", "N", 0),
                                       (256, 256, 64, "N", "T", 1)]:
            A = oracle.uniform(*stored_shape(ta, m, k), 1)
            B = oracle.uniform(*stored_shape(tb, k, n), 2)
            C0 = oracle.uniform(m, n, 3)
            got = run_gpu(handle, ta, tb, m, n, k, 1.0, A, B, 1.0, C0, lda=A.shape[0] + pad)
            assert oracle.rel_fro(got, dd_ref(ta, tb, m, n, k, 1.0, A, B, 1.0, C0)) <= gemm_bound(k)
    finally:
        ca.set_dgemm_tuning(0, 0)


def test_padded_leading_dimensions_and_unaligned_pointers(handle):
    import torch
    m, n, k = 77, 45, 61
    A, B, C0 = oracle.uniform(m, k, 1), oracle.uniform(k, n, 2), oracle.uniform(m, n, 3)
    # padded leading dimensions (odd ones disable 16-byte loads)
    got = run_gpu(handle, "N", "N", m, n, k, 1.0, A, B, 0.5, C0, lda=m + 3, ldb=k + 5, ldc=m + 1)
    ref = dd_ref("N", "N", m, n, k, 1.0, A, B, 0.5, C0)
    assert oracle.rel_fro(got, ref) <= gemm_bound(k)
    # operands starting one double past a 16-byte boundary
    dA, _ = to_dev(np.vstack([np.zeros((1, k)), A]).copy(order="F"))
    dB, _ = to_dev(np.vstack([np.zeros((1, n)), B]).copy(order="F"))
    dC, _ = to_dev(np.vstack([np.zeros((1, n)), C0]).copy(order="F"))
    handle.dgemm("N", "N", m, n, k, 1.0, dA.data_ptr() + 8, m + 1, dB.data_ptr() + 8, k + 1, 0.5,
                 dC.data_ptr() + 8, m + 1)
    torch.cuda.synchronize()
    got = from_dev(dC, m + 1)[1:, :]
    assert oracle.rel_fro(got, ref) <= gemm_bound(k)


def test_beta_zero_never_reads_c(handle):
    m, n, k = 200, 150, 40
    A, B = oracle.uniform(m, k, 1), oracle.uniform(k, n, 2)
    got = run_gpu(handle, "N", "N", m, n, k, 1.0, A, B, 0.0, np.zeros((m, n), order="F"), nan_c=True)
    assert np.isfinite(got).all()


def test_degenerate_extents_and_alpha_zero(handle):
    import torch
    m, n, k = 50, 40, 30
    A, B, C0 = oracle.uniform(m, k, 1), oracle.uniform(k, n, 2), oracle.uniform(m, n, 3)
    for kk, alpha, beta in ((0, 1.0, 0.5), (k, 0.0, 0.5), (0, 1.0, 0.0), (k, 0.0, 1.0)):
        got = run_gpu(handle, "N", "N", m, n, kk, alpha, A, B, beta, C0)
        assert np.array_equal(got, beta * C0 if beta != 0 else np.zeros_like(C0)), (kk, alpha, beta)
    dC, ldc = to_dev(C0)
    for mm, nn in ((0, n), (m, 0)):
        handle.dgemm("N", "N", mm, nn, k, 1.0, 0, m, 0, k, 0.0, dC.data_ptr(), ldc)
    torch.cuda.synchronize()
    assert np.array_equal(from_dev(dC, m), C0)


def test_invalid_arguments_raise(handle):
    import cosma_amd as ca
    with pytest.raises(ca.CosmaAmdError, match="transa"):
        handle.dgemm("X", "N", 4, 4, 4, 1.0, 16, 4, 16, 4, 0.0, 16, 4)
    with pytest.raises(ca.CosmaAmdError, match="lda"):
        handle.dgemm("N", "N", 8, 4, 4, 1.0, 16, 4, 16, 4, 0.0, 16, 8)
    with pytest.raises(ca.CosmaAmdError, match="negative"):
        handle.dgemm("N", "N", -1, 4, 4, 1.0, 16, 4, 16, 4, 0.0, 16, 4)


def test_miniapp_shape_2000_against_host_blas(handle):
    # cosma_miniapp -m 2000 -n 2000 -k 2000 on one rank: one NN dgemm, alpha=1, beta=0.
    m = n = k = 2000
    A, B = oracle.uniform(m, k, 21), oracle.uniform(k, n, 22)
    got = run_gpu(handle, "N", "N", m, n, k, 1.0, A, B, 0.0, np.zeros((m, n), order="F"), nan_c=True)
    ref = np.zeros((m, n), order="F")
    oracle.hostblas.dgemm("N", "N", m, n, k, 1.0, A, m, B, k, 0.0, ref, m)
    assert oracle.rel_fro(got, ref) <= gemm_bound(k)


def test_fill_uniform_matches_host_generator(handle):
    import torch
    import cosma_amd as ca
    rows, cols, ld = 1000, 37, 1003
    t = torch.full((cols, ld), 7.0, dtype=torch.float64, device="cuda")
    ca.fill_uniform(t.data_ptr(), rows, cols, ld, 1234, 0)
    torch.cuda.synchronize()
    got = from_dev(t, rows)
    assert np.array_equal(got, oracle.uniform(rows, cols, 1234))
    assert (t[:, rows:].cpu() == 7.0).all()


def test_full_size_16384_tn_sampled_rows(handle):
    """Config-2 shape on device-resident operands: m = n = k = 16384,
    op(A) = A^T, alpha = 1, beta = 0.5.  The whole of C must be finite and
    sampled rows of C must match the double-double oracle within the FP64
    bound."""
    import torch
    import cosma_amd as ca
    N = 16384
    A = torch.empty((N, N), dtype=torch.float64, device="cuda")   # stored k x m, ld = N
    B = torch.empty((N, N), dtype=torch.float64, device="cuda")
    C = torch.empty((N, N), dtype=torch.float64, device="cuda")
    ca.fill_uniform(A.data_ptr(), N, N, N, 11)
    ca.fill_uniform(B.data_ptr(), N, N, N, 12)
    ca.fill_uniform(C.data_ptr(), N, N, N, 13)
    handle.dgemm("T", "N", N, N, N, 1.0, A.data_ptr(), N, B.data_ptr(), N, 0.5, C.data_ptr(), N)
    torch.cuda.synchronize()
    assert torch.isfinite(C).all().item()
    Bh = np.asfortranarray(B.cpu().numpy().T)      # column-major k x n
    for i in (0, 12345):
        a_row = np.asfortranarray(A[i].cpu().numpy().reshape(N, 1))   # op(A)(i, :) = stored column i
        ref = np.zeros((1, N), order="F")
        oracle.dgemm("T", "N", 1, N, N, 1.0, a_row, N, Bh, N, 0.0, ref, 1, dd=True)
        c0 = oracle.uniform_at(13, i + np.arange(N, dtype=np.uint64) * N)   # row i of C0
        got = C[:, i].cpu().numpy()                                         # row i of C
        assert oracle.rel_fro(got, ref[0] + 0.5 * c0) <= gemm_bound(N), i
