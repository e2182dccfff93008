"""Grouped (batched) DGEMM plans through the C ABI.

Each problem of a plan must come out bit-identical to cosma_amd_dgemm on the
same operands (the grouped kernel runs the same tile routine), for every tile
configuration, for the direct launch and for the captured hipGraph, and the
graph must give the eager result on every replay.  A sample of problems is
also checked against the double-double oracle to the GEMM bound.
"""
import numpy as np
import pytest

import oracle
from helpers import from_dev, gemm_bound, stored_shape, to_dev

pytestmark = pytest.mark.gpu


def make_problems(count, seed, max_dim=200):
    import torch
    rng = np.random.default_rng(seed)
    probs = []
    for i in range(count):
        m, n, k = (int(x) for x in rng.integers(1, max_dim + 1, 3))
        if i % 17 == 3:
            k = 0
        if i % 19 == 5:
            m = 0
        ta, tb = "NT"[int(rng.integers(0, 2))], "NT"[int(rng.integers(0, 2))]
        alpha = float(rng.choice([1.0, -1.3, 0.0 if i % 23 == 7 else 0.5]))
        beta = float(rng.choice([0.0, 0.5, 1.0]))
        A = oracle.uniform(*stored_shape(ta, max(m, 1), max(k, 1)), 1000 + i)
        B = oracle.uniform(*stored_shape(tb, max(k, 1), max(n, 1)), 2000 + i)
        C0 = oracle.uniform(max(m, 1), max(n, 1), 3000 + i)
        pad = int(rng.integers(0, 2))
        dA, lda = to_dev(A, A.shape[0] + pad)
        dB, ldb = to_dev(B, B.shape[0] + pad)
        dC, ldc = to_dev(C0)
        if beta == 0.0:
            dC.fill_(float("nan"))
        probs.append(dict(ta=ta, tb=tb, m=m, n=n, k=k, alpha=alpha, beta=beta, A=A, B=B, C0=C0,
                          dA=dA, lda=lda, dB=dB, ldb=ldb, dC=dC, ldc=ldc))
    torch.cuda.synchronize()
    return probs


def descs(probs, key="dC"):
    return [(p["ta"], p["tb"], p["m"], p["n"], p["k"], p["alpha"], p["dA"].data_ptr(), p["lda"],
             p["dB"].data_ptr(), p["ldb"], p["beta"], p[key].data_ptr(), p["ldc"]) for p in probs]


def reset_c(probs):
    for p in probs:
        src, _ = to_dev(p["C0"])
        p["dC"].copy_(src)
        if p["beta"] == 0.0:
            p["dC"].fill_(float("nan"))


def single_results(handle, probs):
    """The same problems through cosma_amd_dgemm, one call each."""
    import torch
    out = []
    for p in probs:
        c = p["dC"].clone()
        handle.dgemm(p["ta"], p["tb"], p["m"], p["n"], p["k"], p["alpha"], p["dA"].data_ptr(), p["lda"],
                     p["dB"].data_ptr(), p["ldb"], p["beta"], c.data_ptr(), p["ldc"])
        out.append(c)
    torch.cuda.synchronize()
    return [from_dev(c, max(p["m"], 1)) for c, p in zip(out, probs)]


def grouped_results(probs):
    return [from_dev(p["dC"], max(p["m"], 1)) for p in probs]


def same(got, exp):
    return all(np.array_equal(g, e, equal_nan=True) for g, e in zip(got, exp))


@pytest.fixture(scope="module")
def problems(handle):
    probs = make_problems(60, 11)
    return probs, single_results(handle, probs)


@pytest.mark.parametrize("cfg", [1, 2, 3, 4, 5])
def test_grouped_run_bit_identical_to_single_calls(handle, problems, cfg):
    import torch
    import cosma_amd as ca
    probs, expected = problems
    reset_c(probs)
    plan = ca.Grouped(handle, descs(probs), tile_config=cfg)
    assert plan.info()["tile_config"] == cfg
    plan.run()
    torch.cuda.synchronize()
    assert same(grouped_results(probs), expected)
    plan.close()


def test_grouped_against_oracle(handle, problems):
    import torch
    import cosma_amd as ca
    probs, _ = problems
    reset_c(probs)
    plan = ca.Grouped(handle, descs(probs), tile_config=1)
    plan.run()
    torch.cuda.synchronize()
    for p, got in zip(probs, grouped_results(probs)):
        m, n, k = p["m"], p["n"], p["k"]
        if m == 0 or n == 0:
            continue
        ref = p["C0"].copy(order="F")
        oracle.dgemm(p["ta"], p["tb"], m, n, k, p["alpha"], p["A"], p["A"].shape[0], p["B"], p["B"].shape[0],
                     p["beta"], ref, m, dd=True)
        if k == 0 or p["alpha"] == 0.0:
            assert np.array_equal(got, p["beta"] * p["C0"] if p["beta"] != 0 else np.zeros_like(ref))
        else:
            assert oracle.rel_fro(got, ref) <= gemm_bound(k), (m, n, k)


def test_graph_replays_match_eager(handle, problems):
    """Replays with eager work (resetting C) between them; every replay must
    equal the eager result (beta != 0 problems read the reset C each time)."""
    import torch
    import cosma_amd as ca
    probs, expected = problems
    plan = ca.Grouped(handle, descs(probs), tile_config=2)
    for rep in range(3):
        reset_c(probs)
        torch.cuda.synchronize()
        plan.graph_run()
        torch.cuda.synchronize()
        assert same(grouped_results(probs), expected), f"replay {rep}"
    plan.close()


def test_tuner_picks_a_valid_configuration(handle, problems):
    import torch
    import cosma_amd as ca
    probs, expected = problems
    plan = ca.Grouped(handle, descs(probs))
    plan.tune()
    info = plan.info()
    assert info["tile_config"] in (1, 2, 3, 4, 5) and 1 <= info["grid"] <= max(1, info["tiles"])
    # a second plan over the same shapes starts from the tuned choice
    plan2 = ca.Grouped(handle, descs(probs))
    assert plan2.info()["tile_config"] == info["tile_config"] and plan2.info()["grid"] == info["grid"]
    reset_c(probs)
    plan.graph_run()
    torch.cuda.synchronize()
    assert same(grouped_results(probs), expected)
    plan.close()
    plan2.close()


def test_invalid_problems_are_rejected(handle):
    import torch
    import cosma_amd as ca
    a = torch.zeros((8, 8), dtype=torch.float64, device="cuda")
    host = np.zeros((8, 8))
    good = ("N", "N", 8, 8, 8, 1.0, a.data_ptr(), 8, a.data_ptr(), 8, 0.0, a.data_ptr(), 8)
    with pytest.raises(ca.CosmaAmdError, match="device memory"):
        ca.Grouped(handle, [good, ("N", "N", 8, 8, 8, 1.0, host.ctypes.data, 8, a.data_ptr(), 8, 0.0,
                                   a.data_ptr(), 8)])
    with pytest.raises(ca.CosmaAmdError, match="leading dimension"):
        ca.Grouped(handle, [("N", "N", 8, 8, 8, 1.0, a.data_ptr(), 4, a.data_ptr(), 8, 0.0, a.data_ptr(), 8)])
    with pytest.raises(ca.CosmaAmdError, match="transpose"):
        ca.Grouped(handle, [("X", "N", 8, 8, 8, 1.0, a.data_ptr(), 8, a.data_ptr(), 8, 0.0, a.data_ptr(), 8)])
    empty = ca.Grouped(handle, [])
    empty.run()
    empty.graph_run()
    empty.close()
