"""Host-streaming DGEMM (the Tiled-MM replacement) against the oracle.

Covers the memory kinds the streamer accepts (device, pinned host, pageable
host, mixed), both schedules (resident panels; bounded ring forced with a
small memory budget), small tiles as the reference's GPU CI uses
(COSMA_GPU_MAX_TILE_* = 100, ci/cscs.yml:45-49), beta == 0 with NaN in host C,
k == 0 / alpha == 0, copy_c_back = false, and the k-panel streaming shape of
the tall-skinny configuration at reduced size."""
import numpy as np
import pytest

import oracle
from helpers import gemm_bound, stored_shape

pytestmark = pytest.mark.gpu


def make_handle(tile, streams=2):
    import cosma_amd as ca
    return ca.Handle(0, streams=streams, tile_m=tile[0], tile_n=tile[1], tile_k=tile[2])


def host_buf(a, kind):
    """Column-major numpy matrix -> (owner, pointer, ld) in the requested memory kind."""
    import torch
    a = np.asfortranarray(a)
    t = torch.from_numpy(np.array(a.T, order="C"))           # (cols, rows), always a copy
    if kind == "pinned":
        t = t.pin_memory()
    elif kind == "device":
        t = t.cuda()
    return t, t.data_ptr(), a.shape[0]


def read_back(t, rows):
    return np.asfortranarray(t[:, :rows].cpu().numpy().T)


def run(h, ta, tb, m, n, k, alpha, beta, kinds, seed=0, nan_c=False, copy_c_back=True):
    import torch
    A = oracle.uniform(*stored_shape(ta, m, k), seed + 1)
    B = oracle.uniform(*stored_shape(tb, k, n), seed + 2)
    C0 = oracle.uniform(m, n, seed + 3)
    ref = C0.copy(order="F")
    oracle.dgemm(ta, tb, m, n, k, alpha, A, A.shape[0], B, B.shape[0], beta, ref, m, dd=True)
    tA, pA, lda = host_buf(A, kinds[0])
    tB, pB, ldb = host_buf(B, kinds[1])
    tC, pC, ldc = host_buf(np.full((m, n), np.nan, order="F") if nan_c else C0, kinds[2])
    h.dgemm_host(ta, tb, m, n, k, alpha, pA, lda, pB, ldb, beta, pC, ldc, copy_c_back=copy_c_back)
    torch.cuda.synchronize()
    return tC, ref


KINDS = [("pageable", "pageable", "pageable"), ("pinned", "pinned", "pinned"),
         ("device", "pageable", "pinned"), ("pinned", "device", "pageable"), ("device", "device", "pinned")]


@pytest.mark.parametrize("kinds", KINDS)
@pytest.mark.parametrize("ta,tb", [("N", "N"), ("T", "N"), ("N", "T")])
def test_memory_kinds_resident(kinds, ta, tb):
    h = make_handle((100, 100, 100))
    m, n, k = 257, 190, 333
    tC, ref = run(h, ta, tb, m, n, k, -1.3, 0.5, kinds)
    assert h.last_mode() == 1
    assert oracle.rel_fro(read_back(tC, m), ref) <= gemm_bound(k)


@pytest.mark.parametrize("kinds", KINDS)
@pytest.mark.parametrize("ta,tb,beta", [("T", "N", 0.7), ("N", "T", 1.0), ("N", "N", 0.0)])
def test_resident_schedule_phases(kinds, ta, tb, beta):
    """Three k-panels (one panel-major, two per tile), 3 x 3 uneven tiles (more
    than the 4 deferred-C slots), and a first product split into two 256-deep
    k-chunks; host C is added as beta * C after the last panel, device C takes
    beta in the first product."""
    h = make_handle((256, 192, 512))
    m, n, k = 700, 450, 1300
    tC, ref = run(h, ta, tb, m, n, k, -0.9, beta, kinds, seed=30, nan_c=(beta == 0.0))
    assert h.last_mode() == 1
    assert oracle.rel_fro(read_back(tC, m), ref) <= gemm_bound(k)


@pytest.mark.parametrize("kinds", KINDS[:3])
def test_bounded_ring_schedule(kinds):
    h = make_handle((128, 96, 64), streams=3)
    h.set_memory_limit(1 << 20)               # far below what resident mode needs
    m, n, k = 300, 250, 400
    tC, ref = run(h, "T", "N", m, n, k, 1.0, 1.0, kinds, seed=10)
    assert h.last_mode() == 2
    assert oracle.rel_fro(read_back(tC, m), ref) <= gemm_bound(k)


@pytest.mark.parametrize("kinds", KINDS[:2])
def test_beta_zero_does_not_read_host_c(kinds):
    h = make_handle((64, 64, 64))
    m, n, k = 150, 130, 70
    tC, ref = run(h, "N", "N", m, n, k, 1.0, 0.0, kinds, nan_c=True)
    got = read_back(tC, m)
    assert np.isfinite(got).all()
    assert oracle.rel_fro(got, ref) <= gemm_bound(k)


def test_k_zero_and_alpha_zero_scale_host_c():
    import torch
    h = make_handle((64, 64, 64))
    m, n = 100, 80
    C0 = oracle.uniform(m, n, 5)
    for k, alpha, beta in ((0, 1.0, 0.5), (30, 0.0, -2.0), (0, 1.0, 0.0)):
        tC, pC, ldc = host_buf(C0, "pageable")
        h.dgemm_host("N", "N", m, n, k, alpha, 0, m, 0, max(k, 1), beta, pC, ldc)
        torch.cuda.synchronize()
        assert np.array_equal(read_back(tC, m), beta * C0 if beta != 0 else np.zeros_like(C0))


def test_copy_c_back_false_leaves_result_on_device():
    import torch
    h = make_handle((128, 128, 128))
    m, n, k = 200, 170, 90
    tC, ref = run(h, "N", "N", m, n, k, 1.0, 0.0, ("pinned", "pinned", "pageable"), copy_c_back=False)
    ptr, ld = h.device_c()
    assert ptr and ld == m
    out = torch.empty((n, ld), dtype=torch.float64, device="cuda")
    # copy the device result with the library's own block transform
    h.block_transform("N", m, n, 1.0, ptr, ld, 0.0, out.data_ptr(), ld)
    torch.cuda.synchronize()
    assert oracle.rel_fro(read_back(out, m), ref) <= gemm_bound(k)


def test_tall_skinny_k_panels():
    """Shape of the tall-skinny configuration (A^T B, alpha = beta = 1, C
    pre-filled, k streamed in panels) at 1/8 scale in every dimension:
    m = n = 256, k = 131072, 16 k-panels of 8192."""
    h = make_handle((4096, 4096, 8192))
    m = n = 256
    k = 131072
    tC, ref = run(h, "T", "N", m, n, k, 1.0, 1.0, ("pinned", "pinned", "pinned"), seed=20)
    assert h.last_mode() == 1
    assert oracle.rel_fro(read_back(tC, m), ref) <= gemm_bound(k)


def test_all_device_operands_take_the_direct_path():
    h = make_handle((64, 64, 64))
    m, n, k = 120, 110, 100
    tC, ref = run(h, "N", "T", m, n, k, 2.0, 0.25, ("device", "device", "device"))
    assert h.last_mode() == 0
    assert oracle.rel_fro(read_back(tC, m), ref) <= gemm_bound(k)
