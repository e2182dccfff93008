"""The no-allocation split-precision entry point against the allocating one.

cosma_amd_dgemm_ozaki_noalloc computes in the library's per-device scratch
array instead of the handle's workspaces; its results must be bit-identical
to cosma_amd_dgemm_ozaki's (same statistics, count, bits, residues and
recombination), on two calls in a row whose second is smaller (the scratch
keeps the first call's contents), for the moduli method with INT8 and BF16
residues and for slices, with beta = 0 and beta != 0.  A call too large for
the scratch is refused before any work is queued.
"""
import numpy as np
import pytest

import oracle
from helpers import from_dev, to_dev

pytestmark = pytest.mark.gpu


def call_both(handle, ta, tb, m, n, k, alpha, beta, method, seed, count=0):
    import torch
    A = oracle.uniform(*((m, k) if ta == "N" else (k, m)), seed)
    B = oracle.uniform(*((k, n) if tb == "N" else (n, k)), seed + 1)
    C0 = oracle.uniform(m, n, seed + 2)
    dA, lda = to_dev(A)
    dB, ldb = to_dev(B)
    c1, ldc = to_dev(C0)
    c2, _ = to_dev(C0)
    if beta == 0.0:
        c1.fill_(float("nan"))
        c2.fill_(float("nan"))
    print(f"no-alloc vs allocating: {ta}{tb} m={m} n={n} k={k} beta={beta} method={method} count={count}",
          flush=True)
    i1 = handle.dgemm_ozaki(ta, tb, m, n, k, alpha, dA.data_ptr(), lda, dB.data_ptr(), ldb, beta, c1.data_ptr(),
                            ldc, method=method, count=count)
    i2 = handle.dgemm_ozaki_noalloc(ta, tb, m, n, k, alpha, dA.data_ptr(), lda, dB.data_ptr(), ldb, beta,
                                    c2.data_ptr(), ldc, method=method, count=count)
    torch.cuda.synchronize()
    return from_dev(c1, m), from_dev(c2, m), i1, i2


@pytest.mark.parametrize("method", [0, 2, 1])
def test_noalloc_matches_allocating_call_twice(handle, method):
    for (ta, tb, m, n, k, beta, seed) in (("T", "N", 768, 700, 768, 0.5, 10), ("N", "T", 333, 129, 257, 0.0, 20)):
        got, exp, info_a, info_n = call_both(handle, ta, tb, m, n, k, -0.75, beta, method, seed)
        assert np.array_equal(got, exp)
        for key in ("method", "count", "bits", "int8_products", "tolerance_met", "passes"):
            assert info_a[key] == info_n[key], (key, info_a, info_n)
        assert info_n["product_ms"] == 0.0


def test_noalloc_forced_count_and_cancellation(handle):
    """A forced count, then operands that make the check raise the count (a
    second pass in the same scratch)."""
    got, exp, ia, ina = call_both(handle, "N", "N", 512, 512, 512, 1.0, 0.0, 0, 40, count=20)
    assert np.array_equal(got, exp) and ina["count"] == 20
    import torch
    m = n = 256
    h = 256
    Y, Z, W = oracle.uniform(m, h, 51), oracle.uniform(h, n, 52), oracle.uniform(h, n, 53)
    A = np.asfortranarray(np.hstack([Y, Y]))
    B = np.asfortranarray(np.vstack([Z, -Z + 2.0 ** -8 * W]))
    dA, lda = to_dev(A)
    dB, ldb = to_dev(B)
    c1, ldc = to_dev(np.zeros((m, n), order="F"))
    c2, _ = to_dev(np.zeros((m, n), order="F"))
    i1 = handle.dgemm_ozaki("N", "N", m, n, 2 * h, 1.0, dA.data_ptr(), lda, dB.data_ptr(), ldb, 0.0,
                            c1.data_ptr(), ldc)
    i2 = handle.dgemm_ozaki_noalloc("N", "N", m, n, 2 * h, 1.0, dA.data_ptr(), lda, dB.data_ptr(), ldb, 0.0,
                                    c2.data_ptr(), ldc)
    torch.cuda.synchronize()
    assert i1["passes"] >= 2 and i2["passes"] == i1["passes"] and i2["count"] == i1["count"]
    assert np.array_equal(from_dev(c1, m), from_dev(c2, m))


def test_noalloc_refuses_calls_beyond_the_scratch(handle):
    """m = n = k = 4096 at up to 20 moduli needs about 1 GB: refused before any
    work (the operand pointers are never read)."""
    import torch
    import cosma_amd as ca
    dummy = torch.zeros(16, dtype=torch.float64, device="cuda")
    p = dummy.data_ptr()
    with pytest.raises(ca.CosmaAmdError, match="no allocation"):
        handle.dgemm_ozaki_noalloc("N", "N", 4096, 4096, 4096, 1.0, p, 4096, p, 4096, 0.0, p, 4096)
    torch.cuda.synchronize()
