"""Allocation failures on every entry point that allocates.

Every device and page-locked allocation the library makes goes through one
wrapper, which cosma_amd_test_fail_allocations(skip, count) arms to let
`skip` allocations through and fail the `count` after them.  For each entry
point that allocates (page-locked host buffers, host-tile streaming, the
split-precision GEMM and its grouped plans, the FP64 grouped plans), the tests
fail each of its allocations in turn and check that the call returns
COSMA_AMD_ERR_ALLOC with a message, that nothing is left half-made (a handle
whose workspace failed to grow, or a plan that failed to build), and that the
same call succeeds once the failure is disarmed, with results bit-identical
to a run that never failed.  The reference's memory pool re-throws
std::bad_alloc (src/cosma/memory_pool.cpp:65-89,130-150); the status code is
the C ABI's equivalent."""
import ctypes

import numpy as np
import pytest

import oracle
from helpers import from_dev, gemm_bound, stored_shape, to_dev

pytestmark = pytest.mark.gpu

ERR_ALLOC = 3
MAX_SKIP = 40


@pytest.fixture(autouse=True)
def disarmed():
    import cosma_amd as ca
    ca.fail_allocations(0, 0)
    yield
    ca.fail_allocations(0, 0)


def fail_each_allocation(call):
    """Runs call() with its first, second, ... allocation failing until a run makes no
    allocation past the armed one; returns how many runs failed with ERR_ALLOC."""
    import torch
    import cosma_amd as ca
    failures = 0
    for skip in range(MAX_SKIP):
        ca.fail_allocations(skip, 1)
        try:
            call(skip)
        except ca.CosmaAmdError as e:
            assert e.status == ERR_ALLOC, f"skip {skip}: status {e.status} ({e})"
            assert "fail" in str(e).lower(), f"skip {skip}: no message ({e})"
            failures += 1
            continue
        finally:
            torch.cuda.synchronize()
            ca.fail_allocations(0, 0)
        return failures
    raise AssertionError(f"every one of {MAX_SKIP} runs hit the armed failure")


def test_host_alloc_failure():
    import cosma_amd as ca
    lib = ca.lib()
    p = ctypes.c_void_p()
    ca.fail_allocations(0, 1)
    assert lib.cosma_amd_host_alloc(1 << 20, ctypes.byref(p)) == ERR_ALLOC
    assert not p.value
    assert "host_alloc" in lib.cosma_amd_last_error().decode()
    # the failure was used up: the next allocation goes through
    assert lib.cosma_amd_host_alloc(1 << 20, ctypes.byref(p)) == 0 and p.value
    assert lib.cosma_amd_host_free(p) == 0


def test_count_fails_consecutive_allocations():
    import cosma_amd as ca
    lib = ca.lib()
    p = ctypes.c_void_p()
    ca.fail_allocations(1, 2)
    assert lib.cosma_amd_host_alloc(4096, ctypes.byref(p)) == 0
    assert lib.cosma_amd_host_free(p) == 0
    for _ in range(2):
        assert lib.cosma_amd_host_alloc(4096, ctypes.byref(p)) == ERR_ALLOC
    assert lib.cosma_amd_host_alloc(4096, ctypes.byref(p)) == 0
    assert lib.cosma_amd_host_free(p) == 0


def streaming_case(ta, tb, m, n, k, alpha, beta, seed):
    import torch
    A = oracle.uniform(*stored_shape(ta, m, k), seed + 1)
    B = oracle.uniform(*stored_shape(tb, k, n), seed + 2)
    C0 = oracle.uniform(m, n, seed + 3)
    ref = C0.copy(order="F")
    oracle.dgemm(ta, tb, m, n, k, alpha, A, A.shape[0], B, B.shape[0], beta, ref, m, dd=True)
    host = [torch.from_numpy(np.array(x.T, order="C")) for x in (A, B, C0)]     # pageable
    return host, ref


@pytest.mark.parametrize("budget", [0, 1 << 20])
def test_streaming_allocation_failures(budget):
    """Device tile buffers and the page-locked staging ring of the host-streaming GEMM;
    budget 1 MiB forces the bounded ring schedule."""
    import cosma_amd as ca
    ta, tb, m, n, k, alpha, beta = "N", "T", 300, 260, 340, 1.25, 0.5
    made = []

    def call(skip):
        h = ca.Handle(0, streams=2, tile_m=100, tile_n=100, tile_k=100)
        if budget:
            h.set_memory_limit(budget)
        made.append(h)
        host, _ = streaming_case(ta, tb, m, n, k, alpha, beta, seed=7)
        h.dgemm_host(ta, tb, m, n, k, alpha, host[0].data_ptr(), m, host[1].data_ptr(), n, beta,
                     host[2].data_ptr(), m)

    failures = fail_each_allocation(call)
    assert failures >= 2, f"only {failures} allocation(s) seen"
    # every handle whose call failed must now give the right answer
    results = []
    for h in made:
        host, ref = streaming_case(ta, tb, m, n, k, alpha, beta, seed=7)
        h.dgemm_host(ta, tb, m, n, k, alpha, host[0].data_ptr(), m, host[1].data_ptr(), n, beta,
                     host[2].data_ptr(), m)
        got = np.asfortranarray(host[2].numpy().T)
        err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
        assert err <= gemm_bound(k), err
        results.append(got)
        h.close()
    for got in results[1:]:
        assert np.array_equal(got, results[0])


@pytest.mark.parametrize("method", ["moduli", "slices"])
def test_split_precision_allocation_failures(method):
    """Workspace, auxiliary buffer and page-locked statistics of cosma_amd_dgemm_ozaki."""
    import torch
    import cosma_amd as ca
    m, n, k = 320, 288, 352
    A = oracle.uniform(m, k, 11)
    B = oracle.uniform(k, n, 12)
    C0 = oracle.uniform(m, n, 13)
    dA, lda = to_dev(A)
    dB, ldb = to_dev(B)
    code = ca.OZAKI_MODULI if method == "moduli" else ca.OZAKI_SLICES
    made = []

    def run(h):
        dC, ldc = to_dev(C0)
        info = h.dgemm_ozaki("N", "N", m, n, k, 1.0, dA.data_ptr(), lda, dB.data_ptr(), ldb, 0.75,
                             dC.data_ptr(), ldc, method=code)
        torch.cuda.synchronize()
        return from_dev(dC, m), info

    def call(skip):
        h = ca.Handle(0)
        made.append(h)
        run(h)

    failures = fail_each_allocation(call)
    assert failures >= 2, f"only {failures} allocation(s) seen"
    fresh = ca.Handle(0)
    want, want_info = run(fresh)
    fresh.close()
    for h in made:
        got, info = run(h)
        assert np.array_equal(got, want)
        assert info["count"] == want_info["count"]
        h.close()


def grouped_problems(count, seed, max_dim, big=()):
    import torch
    rng = np.random.default_rng(seed)
    probs = []
    dims = [tuple(int(x) for x in rng.integers(1, max_dim + 1, 3)) for _ in range(count)] + list(big)
    for i, (m, n, k) in enumerate(dims):
        A = oracle.uniform(m, k, 500 + i)
        B = oracle.uniform(n, k, 600 + i)
        C0 = oracle.uniform(m, n, 700 + i)
        dA, lda = to_dev(A)
        dB, ldb = to_dev(B)
        dC, ldc = to_dev(C0)
        probs.append(dict(m=m, n=n, k=k, C0=C0, dA=dA, lda=lda, dB=dB, ldb=ldb, dC=dC, ldc=ldc))
    torch.cuda.synchronize()
    return probs


def grouped_descs(probs):
    return [("N", "T", p["m"], p["n"], p["k"], 1.5, p["dA"].data_ptr(), p["lda"], p["dB"].data_ptr(),
             p["ldb"], 0.5, p["dC"].data_ptr(), p["ldc"]) for p in probs]


def grouped_results(probs, plan, graph=False):
    import torch
    for p in probs:
        p["dC"].copy_(to_dev(p["C0"], p["ldc"])[0])
    torch.cuda.synchronize()
    plan.graph_run() if graph else plan.run()
    torch.cuda.synchronize()
    return [from_dev(p["dC"], p["m"]) for p in probs]


def test_grouped_plan_allocation_failures():
    """Descriptor arrays of the FP64 grouped plan (one per tile configuration)."""
    import cosma_amd as ca
    probs = grouped_problems(40, 3, 150)
    h = ca.Handle(0)

    def call(skip):
        ca.Grouped(h, grouped_descs(probs)).close()

    failures = fail_each_allocation(call)
    assert failures >= 1
    plan = ca.Grouped(h, grouped_descs(probs))
    want = grouped_results(probs, plan)
    for p, got in zip(probs, want):
        ref = p["C0"].copy(order="F")
        oracle.dgemm("N", "T", p["m"], p["n"], p["k"], 1.5, oracle_operand(p, "A"), p["m"],
                     oracle_operand(p, "B"), p["n"], 0.5, ref, p["m"], dd=True)
        assert np.linalg.norm(got - ref) <= gemm_bound(p["k"]) * max(np.linalg.norm(ref), 1e-300)
    assert all(np.array_equal(a, b) for a, b in zip(grouped_results(probs, plan, graph=True), want))
    plan.close()
    h.close()


def oracle_operand(p, which):
    t = p["dA"] if which == "A" else p["dB"]
    rows = p["m"] if which == "A" else p["n"]
    return from_dev(t, rows)


@pytest.mark.parametrize("split_min_dim", [0, 64])
def test_grouped_split_plan_allocation_failures(split_min_dim):
    """Buffers of a grouped split-precision plan, and with a threshold of 64 also the FP64
    plan of its small problems; then the plan's error check, which allocates per call."""
    import cosma_amd as ca
    probs = grouped_problems(12, 5, 60, big=[(96, 80, 112), (72, 128, 90)])
    h = ca.Handle(0)

    def call(skip):
        ca.GroupedOzaki(h, grouped_descs(probs), moduli=12, split_min_dim=split_min_dim).close()

    failures = fail_each_allocation(call)
    assert failures >= 2, f"only {failures} allocation(s) seen"
    plan = ca.GroupedOzaki(h, grouped_descs(probs), moduli=12, split_min_dim=split_min_dim)
    info = plan.info()
    if split_min_dim:
        assert info["split_problems"] == 2 and info["fp64_problems"] == 12
    got = grouped_results(probs, plan)
    other = ca.Handle(0)
    clean = ca.GroupedOzaki(other, grouped_descs(probs), moduli=12, split_min_dim=split_min_dim)
    want = grouped_results(probs, clean)
    assert all(np.array_equal(a, b) for a, b in zip(got, want))
    assert all(np.array_equal(a, b) for a, b in zip(grouped_results(probs, plan, graph=True), want))

    ca.fail_allocations(0, 1)
    with pytest.raises(ca.CosmaAmdError) as e:
        plan.check()
    assert e.value.status == ERR_ALLOC
    ca.fail_allocations(0, 0)
    assert plan.check() == clean.check()
    for x in (plan, clean, h, other):
        x.close()
