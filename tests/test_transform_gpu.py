"""Parity of the block transform kernel (COSTA's local reorder/transpose)
with the oracle, through the C ABI.  alpha = 1, beta = 0 is a permutation and
must be bit-exact; other alpha/beta combinations use one fused multiply-add
per element, the same formula as the oracle, so they are bit-exact too."""
import numpy as np
import pytest

import oracle
from helpers import from_dev, load_transform, to_dev, transform_fixtures

pytestmark = pytest.mark.gpu


def run_transform(handle, op, rows, cols, alpha, src, lds, beta, dst0, ldd, nan_dst=False):
    import torch
    dS, _ = to_dev(src)
    dD, _ = to_dev(dst0)
    if nan_dst:
        dD[:, :rows] = float("nan")
    handle.block_transform(op, rows, cols, alpha, dS.data_ptr(), lds, beta, dD.data_ptr(), ldd)
    torch.cuda.synchronize()
    return from_dev(dD, ldd)


@pytest.mark.parametrize("name", transform_fixtures())
def test_golden_fixtures(handle, name):
    d = load_transform(name)
    got = run_transform(handle, d["op"], d["rows"], d["cols"], d["alpha"], d["src"], d["lds"], d["beta"],
                        d["dst0"], d["ldd"])
    assert np.array_equal(got, d["dst"])


@pytest.mark.parametrize("op", "NT")
@pytest.mark.parametrize("rows,cols", [(1, 1), (63, 65), (64, 64), (129, 257), (1000, 3), (3, 1000)])
@pytest.mark.parametrize("alpha,beta", [(1.0, 0.0), (-1.3, 0.5), (2.0, 0.0)])
def test_random_shapes(handle, op, rows, cols, alpha, beta):
    sr, sc = (rows, cols) if op == "N" else (cols, rows)
    lds, ldd = sr + 3, rows + 2
    src, dst0 = oracle.uniform(lds, sc, rows + cols), oracle.uniform(ldd, cols, 7)
    got = run_transform(handle, op, rows, cols, alpha, src, lds, beta, dst0, ldd, nan_dst=(beta == 0.0))
    ref = dst0.copy(order="F")
    if beta == 0.0:
        ref[:rows, :] = np.nan
    oracle.block_transform(op, rows, cols, alpha, src, lds, beta, ref, ldd)
    assert np.array_equal(got[:rows], ref[:rows]), "values"
    assert np.array_equal(got[rows:], dst0[rows:]), "padding rows must be untouched"


def test_scalapack_128_block_reorder_16384(handle):
    """Config-2 reorder: a 16384 x 16384 matrix in ScaLAPACK 128 x 128
    block-cyclic layout on a 1 x 1 process grid (local array = the global
    matrix, lld = 16384) into COSMA's column-major layout, transposed as
    op(A) = A^T requires.  A permutation: must be bit-exact."""
    import torch
    import cosma_amd as ca
    N = 16384
    assert oracle.numroc(N, 128, 0, 0, 1) == N
    S = torch.empty((N, N), dtype=torch.float64, device="cuda")
    D = torch.empty((N, N), dtype=torch.float64, device="cuda")
    ca.fill_uniform(S.data_ptr(), N, N, N, 31)
    handle.block_transform("T", N, N, 1.0, S.data_ptr(), N, 0.0, D.data_ptr(), N)
    torch.cuda.synchronize()
    assert torch.equal(D, S.t())
    handle.block_transform("N", N, N, 1.0, S.data_ptr(), N, 0.0, D.data_ptr(), N)
    torch.cuda.synchronize()
    assert torch.equal(D, S)
