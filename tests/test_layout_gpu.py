"""Grid-layout multiply (the single-rank work of dmultiply_using_layout).

Each matrix is cut by irregular row and column splits into blocks that live
in separate buffers (device, pinned host or pageable host memory, with padded
leading dimensions), as the C API's `layout` struct describes them
(src/cosma/cinterface.hpp:16-40).  The result is checked
  * bit for bit against the same product assembled by hand on the device
    (DGEMM with (1, 0), then the block transform with (alpha, beta)): the
    gather and scatter of blocks must not change a single bit;
  * against the double-double oracle to the GEMM error bound.
"""
import numpy as np
import pytest

import oracle
from helpers import from_dev, gemm_bound, stored_shape, to_dev

pytestmark = pytest.mark.gpu

KINDS = ("device", "pinned", "pageable")


def splits(n, parts, rng):
    if n == 0:
        return [0, 0]
    cuts = sorted(rng.choice(np.arange(1, n), size=min(parts - 1, max(n - 1, 0)), replace=False)) if n > 1 else []
    return [0] + [int(c) for c in cuts] + [n]


class Grid:
    """A matrix split into blocks, each in its own buffer of some memory kind."""

    def __init__(self, X, rsplit, csplit, kinds, rng):
        import torch
        self.X = np.asfortranarray(X)
        self.rsplit, self.csplit = rsplit, csplit
        self.blocks = []      # (tensor, kind, ld, r0, c0, rows, cols)
        i = 0
        for bi in range(len(rsplit) - 1):
            for bj in range(len(csplit) - 1):
                r0, r1, c0, c1 = rsplit[bi], rsplit[bi + 1], csplit[bj], csplit[bj + 1]
                rows, cols = r1 - r0, c1 - c0
                ld = rows + int(rng.integers(0, 3))
                kind = kinds[i % len(kinds)]
                i += 1
                buf = np.zeros((max(cols, 1), max(ld, 1)))
                buf[:cols, :rows] = self.X[r0:r1, c0:c1].T
                t = torch.from_numpy(buf)
                if kind == "pinned":
                    t = t.pin_memory()
                elif kind == "device":
                    t = t.cuda()
                self.blocks.append((t, kind, max(ld, 1), r0, c0, rows, cols))

    def descs(self):
        return [(t.data_ptr(), ld, r0, c0, rows, cols) for (t, _, ld, r0, c0, rows, cols) in self.blocks]

    def gather(self):
        out = np.zeros_like(self.X)
        for (t, _, ld, r0, c0, rows, cols) in self.blocks:
            out[r0:r0 + rows, c0:c0 + cols] = t.cpu().numpy()[:cols, :rows].T
        return out


def make_grid(X, parts, kinds, rng):
    return Grid(X, splits(X.shape[0], parts[0], rng), splits(X.shape[1], parts[1], rng), kinds, rng)


def device_reference(handle, ta, tb, m, n, k, alpha, A, B, beta, C0):
    """Same kernels, same order, on contiguous device copies."""
    import torch
    dA, lda = to_dev(A)
    dB, ldb = to_dev(B)
    dX = torch.zeros((n, m), dtype=torch.float64, device="cuda")
    dC, ldc = to_dev(C0)
    handle.dgemm(ta, tb, m, n, k, 1.0, dA.data_ptr(), lda, dB.data_ptr(), ldb, 0.0, dX.data_ptr(), m)
    handle.block_transform("N", m, n, alpha, dX.data_ptr(), m, beta, dC.data_ptr(), ldc)
    torch.cuda.synchronize()
    return from_dev(dC, m)


@pytest.mark.parametrize("ta,tb", [("N", "N"), ("T", "N"), ("N", "T"), ("T", "T")])
@pytest.mark.parametrize("alpha,beta", [(1.0, 0.0), (-1.3, 0.5), (0.7, 1.0)])
def test_layout_matches_device_product(handle, ta, tb, alpha, beta):
    rng = np.random.default_rng([ord(ta), ord(tb), int(alpha * 10) + 20, int(beta * 10)])
    m, n, k = 157, 94, 213
    A = oracle.uniform(*stored_shape(ta, m, k), 11)
    B = oracle.uniform(*stored_shape(tb, k, n), 12)
    C0 = oracle.uniform(m, n, 13)
    gA = make_grid(A, (3, 4), KINDS, rng)
    gB = make_grid(B, (2, 3), KINDS[::-1], rng)
    gC = make_grid(C0, (4, 2), KINDS, rng)
    handle.dgemm_layout(ta, tb, m, n, k, alpha, gA.descs(), gB.descs(), beta, gC.descs())
    got = gC.gather()
    assert np.array_equal(got, device_reference(handle, ta, tb, m, n, k, alpha, A, B, beta, C0))
    ref = C0.copy(order="F")
    oracle.dgemm(ta, tb, m, n, k, alpha, A, A.shape[0], B, B.shape[0], beta, ref, m, dd=True)
    assert oracle.rel_fro(got, ref) <= gemm_bound(k)


def test_block_cyclic_layout_single_rank(handle):
    """A 1x1-grid ScaLAPACK-style layout: fixed 32 x 48 blocks with ragged edges."""
    rng = np.random.default_rng(5)
    m, n, k = 200, 130, 170
    A, B, C0 = oracle.uniform(m, k, 1), oracle.uniform(k, n, 2), oracle.uniform(m, n, 3)
    def reg(x, b):
        return list(range(0, x, b)) + [x]
    gA = Grid(A, reg(m, 32), reg(k, 48), ("pageable",), rng)
    gB = Grid(B, reg(k, 32), reg(n, 48), ("pageable",), rng)
    gC = Grid(C0, reg(m, 32), reg(n, 48), ("pageable",), rng)
    handle.dgemm_layout("N", "N", m, n, k, 1.0, gA.descs(), gB.descs(), 0.25, gC.descs())
    assert np.array_equal(gC.gather(), device_reference(handle, "N", "N", m, n, k, 1.0, A, B, 0.25, C0))


def test_beta_zero_does_not_read_c_and_scale_only(handle):
    import torch
    rng = np.random.default_rng(7)
    m, n, k = 64, 48, 32
    A, B = oracle.uniform(m, k, 1), oracle.uniform(k, n, 2)
    C0 = np.full((m, n), np.nan, order="F")
    gA, gB = make_grid(A, (2, 2), KINDS, rng), make_grid(B, (2, 2), KINDS, rng)
    gC = make_grid(C0, (3, 2), KINDS, rng)
    handle.dgemm_layout("N", "N", m, n, k, 2.0, gA.descs(), gB.descs(), 0.0, gC.descs())
    got = gC.gather()
    assert np.isfinite(got).all()
    assert np.array_equal(got, device_reference(handle, "N", "N", m, n, k, 2.0, A, B, 0.0, np.zeros((m, n), order="F")))
    # k == 0 and alpha == 0: C <- beta * C (exact), A and B are not touched
    C1 = oracle.uniform(m, n, 4)
    gC = make_grid(C1, (2, 3), KINDS, rng)
    handle.dgemm_layout("N", "N", m, n, 0, 1.0, [], [], 0.5, gC.descs())
    assert np.array_equal(gC.gather(), 0.5 * C1)
    gC = make_grid(C1, (2, 3), KINDS, rng)
    handle.dgemm_layout("N", "N", m, n, k, 0.0, gA.descs(), gB.descs(), 3.0, gC.descs())
    assert np.array_equal(gC.gather(), 3.0 * C1)
    torch.cuda.synchronize()


def test_invalid_tilings_are_rejected(handle):
    import cosma_amd as ca
    import torch
    m = n = k = 8
    buf = torch.zeros((8, 8), dtype=torch.float64, device="cuda")
    p = buf.data_ptr()
    full = [(p, 8, 0, 0, 8, 8)]
    with pytest.raises(ca.CosmaAmdError, match="cover"):          # gap
        handle.dgemm_layout("N", "N", m, n, k, 1.0, [(p, 8, 0, 0, 8, 4)], full, 0.0, full)
    with pytest.raises(ca.CosmaAmdError, match="overlap"):        # overlap (areas add up)
        handle.dgemm_layout("N", "N", m, n, k, 1.0, [(p, 8, 0, 0, 8, 5), (p, 8, 0, 3, 8, 3)], full, 0.0, full)
    with pytest.raises(ca.CosmaAmdError, match="outside"):
        handle.dgemm_layout("N", "N", m, n, k, 1.0, [(p, 8, 1, 0, 8, 8)], full, 0.0, full)
    with pytest.raises(ca.CosmaAmdError, match="ld"):
        handle.dgemm_layout("N", "N", m, n, k, 1.0, [(p, 4, 0, 0, 8, 8)], full, 0.0, full)
