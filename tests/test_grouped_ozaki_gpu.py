"""Grouped split-precision (moduli method) plans through the C ABI.

Problems below a plan's split threshold (min(m, n, k) < 2048 by default) run
on the FP64 grouped kernel and must come out bit-identical to cosma_amd_dgemm
on them; the tests of the split method force it with a threshold of 0.
With the same moduli count, each split problem of a plan must come out
bit-identical to cosma_amd_dgemm_ozaki on it (same exponents, same P, exact
residue products and recombination), for the direct run and for every replay
of the captured hipGraph.  Problems are also checked against the
double-double oracle at the default tolerance, and the edge cases of the
single-problem path are covered: k == 0, alpha == 0, beta == 0 with NaN in C,
zero rows, and Inf / NaN operands (computed by FP64 dot products)."""
import numpy as np
import pytest

import oracle
from helpers import from_dev, stored_shape, to_dev

pytestmark = pytest.mark.gpu


def make_problems(count, seed, max_dim=200, phi=0.0):
    import torch
    import cosma_amd as ca
    rng = np.random.default_rng(seed)
    probs = []
    for i in range(count):
        m, n, k = (int(x) for x in rng.integers(1, max_dim + 1, 3))
        if i % 17 == 3:
            k = 0
        ta, tb = "NT"[int(rng.integers(0, 2))], "NT"[int(rng.integers(0, 2))]
        alpha = float(rng.choice([1.0, -1.3, 0.0 if i % 23 == 7 else 0.5]))
        beta = float(rng.choice([0.0, 0.5, 1.0]))
        A = oracle.uniform(*stored_shape(ta, m, max(k, 1)), 1000 + i)
        B = oracle.uniform(*stored_shape(tb, max(k, 1), n), 2000 + i)
        if phi:
            A = np.asfortranarray(A * np.exp(phi * rng.standard_normal(A.shape)))
            B = np.asfortranarray(B * np.exp(phi * rng.standard_normal(B.shape)))
        if i % 13 == 2:
            A[:] = 0.0                       # zero operand: zero product
        C0 = oracle.uniform(m, n, 3000 + i)
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


def descs(probs):
    return [(p["ta"], p["tb"], p["m"], p["n"], p["k"], p["alpha"], p["dA"].data_ptr(), p["lda"],
             p["dB"].data_ptr(), p["ldb"], p["beta"], p["dC"].data_ptr(), p["ldc"]) for p in probs]


def reset_c(probs):
    for p in probs:
        src, _ = to_dev(p["C0"])
        p["dC"].copy_(src)
        if p["beta"] == 0.0:
            p["dC"].fill_(float("nan"))


def results(probs):
    return [from_dev(p["dC"], p["m"]) for p in probs]


def single_results(handle, probs, moduli):
    """The same problems through cosma_amd_dgemm_ozaki with the same count."""
    import torch
    import cosma_amd as ca
    out = []
    for p in probs:
        c = p["dC"].clone()
        handle.dgemm_ozaki(p["ta"], p["tb"], p["m"], p["n"], p["k"], p["alpha"], p["dA"].data_ptr(), p["lda"],
                           p["dB"].data_ptr(), p["ldb"], p["beta"], c.data_ptr(), p["ldc"],
                           method=ca.OZAKI_MODULI, count=moduli)
        out.append(c)
    torch.cuda.synchronize()
    return [from_dev(c, p["m"]) for c, p in zip(out, probs)]


def single_dgemm_results(handle, probs):
    """The same problems through the plain FP64 cosma_amd_dgemm."""
    import torch
    out = []
    for p in probs:
        c = p["dC"].clone()
        handle.dgemm(p["ta"], p["tb"], p["m"], p["n"], p["k"], p["alpha"], p["dA"].data_ptr(), p["lda"],
                     p["dB"].data_ptr(), p["ldb"], p["beta"], c.data_ptr(), p["ldc"])
        out.append(c)
    torch.cuda.synchronize()
    return [from_dev(c, p["m"]) for c, p in zip(out, probs)]


def same(got, exp):
    return all(np.array_equal(g, e, equal_nan=True) for g, e in zip(got, exp))


def oracle_result(p):
    m, n, k = p["m"], p["n"], p["k"]
    ref = p["C0"].copy(order="F")
    oracle.dgemm(p["ta"], p["tb"], m, n, k, p["alpha"], p["A"], p["A"].shape[0], p["B"], p["B"].shape[0],
                 p["beta"], ref, m, dd=True)
    return ref


@pytest.fixture(scope="module")
def problems():
    return make_problems(48, 5)


@pytest.mark.parametrize("moduli", [8, 14, 17])
def test_bit_identical_to_single_calls(handle, problems, moduli):
    import torch
    import cosma_amd as ca
    reset_c(problems)
    expected = single_results(handle, problems, moduli)
    plan = ca.GroupedOzaki(handle, descs(problems), moduli=moduli, split_min_dim=0)
    assert plan.info()["moduli"] == moduli
    plan.run()
    torch.cuda.synchronize()
    assert same(results(problems), expected)
    plan.close()


def test_default_tolerance_against_oracle(handle):
    import torch
    import cosma_amd as ca
    probs = make_problems(40, 9, max_dim=300, phi=1.0)
    reset_c(probs)
    plan = ca.GroupedOzaki(handle, descs(probs), split_min_dim=0)
    info = plan.info()
    assert 2 <= info["moduli"] <= 20 and info["error_estimate"] <= 1e-14
    plan.run()
    torch.cuda.synchronize()
    for p, got in zip(probs, results(probs)):
        ref = oracle_result(p)
        if p["k"] == 0 or p["alpha"] == 0.0 or not p["A"].any():
            exp = p["beta"] * p["C0"] if p["beta"] != 0 else np.zeros_like(ref)
            assert np.array_equal(got, exp), (p["m"], p["n"], p["k"])
        else:
            assert oracle.rel_fro(got, ref) <= 1e-14, (p["m"], p["n"], p["k"], oracle.rel_fro(got, ref))
    # the check of the run agrees: every problem's estimate meets the tolerance
    chk = plan.check(1e-14)
    assert 1.5 * chk["worst_estimate"] <= 1e-14 and chk["moduli_needed"] <= info["moduli"], (chk, info)
    plan.close()


def test_check_after_operands_change(handle):
    """A plan fixes its count from the operands present at creation.  When
    later operands cancel (A = [Y, Y], B = [Z; -Z + eps W]: X is eps times
    smaller than |A||B| suggests), the check of a run reports the estimate
    against the tolerance and the count that meets it; a plan with that
    count computes those problems to the tolerance."""
    import torch
    import cosma_amd as ca
    probs = make_problems(12, 31, max_dim=120)
    probs = [p for p in probs if p["k"] >= 4 and p["alpha"] != 0.0 and p["A"].any()]
    reset_c(probs)
    plan = ca.GroupedOzaki(handle, descs(probs), split_min_dim=0)
    base = plan.info()["moduli"]
    plan.run()
    torch.cuda.synchronize()
    assert plan.check(1e-14)["moduli_needed"] <= base
    # new operands in the same buffers: cancelling halves, eps = 2^-9
    rng = np.random.default_rng(7)
    for p in probs:
        m, n, k = p["m"], p["n"], p["k"]
        h = k // 2
        Y = oracle.uniform(m, max(h, 1), 500 + m)
        Z = oracle.uniform(max(h, 1), n, 600 + n)
        W = oracle.uniform(max(h, 1), n, 700 + n)
        Aop = np.zeros((m, k))
        Bop = np.zeros((k, n))
        if h > 0:
            Aop[:, :h], Aop[:, h:2 * h] = Y, Y
            Bop[:h], Bop[h:2 * h] = Z, -Z + 2.0 ** -9 * W
        if k % 2:
            Aop[:, -1] = 0.0
        p["A"] = np.asfortranarray(Aop if p["ta"] == "N" else Aop.T)
        p["B"] = np.asfortranarray(Bop if p["tb"] == "N" else Bop.T)
        p["dA"][:, :p["A"].shape[0]].copy_(torch.from_numpy(np.ascontiguousarray(p["A"].T)))
        p["dB"][:, :p["B"].shape[0]].copy_(torch.from_numpy(np.ascontiguousarray(p["B"].T)))
    reset_c(probs)
    plan.run()
    torch.cuda.synchronize()
    chk = plan.check(1e-14)
    assert 1.5 * chk["worst_estimate"] > 1e-14 and chk["moduli_needed"] > base, (chk, base)
    assert 0 <= chk["worst_problem"] < len(probs)
    plan.close()
    assert chk["moduli_needed"] <= 20
    better = ca.GroupedOzaki(handle, descs(probs), moduli=chk["moduli_needed"], split_min_dim=0)
    reset_c(probs)
    better.run()
    torch.cuda.synchronize()
    assert 1.5 * better.check(1e-14)["worst_estimate"] <= 1e-14
    for p, got in zip(probs, results(probs)):
        ref = oracle_result(p)
        X = np.zeros((p["m"], p["n"]), order="F")
        oracle.dgemm(p["ta"], p["tb"], p["m"], p["n"], p["k"], p["alpha"], p["A"], p["A"].shape[0], p["B"],
                     p["B"].shape[0], 0.0, X, p["m"], dd=True)
        err = np.linalg.norm((got - ref).ravel()) / np.linalg.norm(X.ravel())
        assert err <= 1e-14, (p["m"], p["n"], p["k"], err)
    better.close()


def test_norm_bound_bits_match_single_calls(handle):
    """Operands with one entry per row / column at the maximum and the rest
    far below: the plan's device-side bits rule raises P well above the
    entrywise limit, exactly as the single-problem path does, so the results
    stay bit-identical to single calls (and exact products of the rounded
    inputs), also after the operands change between graph replays."""
    import torch
    import cosma_amd as ca
    rng = np.random.default_rng(13)
    probs = make_problems(12, 17, max_dim=160)
    for p in probs:
        if p["k"] == 0:
            continue
        A = p["A"] * 2.0 ** -14
        B = p["B"] * 2.0 ** -14
        rows, cols = A.shape
        if p["ta"] == "N":
            A[np.arange(rows), rng.integers(0, cols, rows)] = 1.0
        else:
            A[rng.integers(0, rows, cols), np.arange(cols)] = 1.0
        rows, cols = B.shape
        if p["tb"] == "N":
            B[rng.integers(0, rows, cols), np.arange(cols)] = -1.0
        else:
            B[np.arange(rows), rng.integers(0, cols, rows)] = -1.0
        p["A"], p["B"] = np.asfortranarray(A), np.asfortranarray(B)
        p["dA"].copy_(to_dev(p["A"], p["lda"])[0])
        p["dB"].copy_(to_dev(p["B"], p["ldb"])[0])
    reset_c(probs)
    torch.cuda.synchronize()
    expected = single_results(handle, probs, 8)
    plan = ca.GroupedOzaki(handle, descs(probs), moduli=8, split_min_dim=0)
    plan.graph_run()
    torch.cuda.synchronize()
    assert same(results(probs), expected)
    # new operands of the same shapes (uniform, no spikes): the replay
    # recomputes the bits from them
    for i, p in enumerate(probs):
        p["A"] = np.asfortranarray(oracle.uniform(*p["A"].shape, 500 + i))
        p["B"] = np.asfortranarray(oracle.uniform(*p["B"].shape, 600 + i))
        p["dA"].copy_(to_dev(p["A"], p["lda"])[0])
        p["dB"].copy_(to_dev(p["B"], p["ldb"])[0])
    reset_c(probs)
    torch.cuda.synchronize()
    expected = single_results(handle, probs, 8)
    plan.graph_run()
    torch.cuda.synchronize()
    assert same(results(probs), expected)
    plan.close()


def test_graph_replays_match_eager(handle, problems):
    import torch
    import cosma_amd as ca
    reset_c(problems)
    plan = ca.GroupedOzaki(handle, descs(problems), moduli=12, split_min_dim=0)
    plan.run()
    torch.cuda.synchronize()
    expected = results(problems)
    for rep in range(3):
        reset_c(problems)
        torch.cuda.synchronize()
        plan.graph_run()
        torch.cuda.synchronize()
        assert same(results(problems), expected), f"replay {rep}"
    plan.close()


def test_tuner_keeps_results(handle, problems):
    import torch
    import cosma_amd as ca
    reset_c(problems)
    plan = ca.GroupedOzaki(handle, descs(problems), moduli=10, split_min_dim=0)
    plan.run()
    torch.cuda.synchronize()
    expected = results(problems)
    plan.tune()
    info = plan.info()
    assert 1 <= info["grid"] <= max(1, info["tiles"])
    reset_c(problems)
    plan.graph_run()
    torch.cuda.synchronize()
    assert same(results(problems), expected)
    plan.set_grid(3)
    reset_c(problems)
    plan.run()
    torch.cuda.synchronize()
    assert same(results(problems), expected)
    plan.close()


def test_non_finite_operands_take_fp64(handle):
    import torch
    import cosma_amd as ca
    probs = make_problems(6, 21, max_dim=90)
    probs = [p for p in probs if p["k"] > 0 and p["alpha"] != 0.0 and p["A"].any()]
    bad = probs[0]
    # an Inf at op(A)(0, 0) (stored entry (0, 0) either way): row 0 of the
    # product is non-finite, the rest finite
    bad["A"][0, 0] = np.inf
    bad["dA"], bad["lda"] = to_dev(bad["A"], bad["lda"])
    reset_c(probs)
    # the non-finite problem does not use the moduli: it does not drive the count
    auto = ca.GroupedOzaki(handle, descs(probs), split_min_dim=0)
    assert auto.info()["moduli"] < 20 and np.isfinite(auto.info()["error_estimate"])
    auto.close()
    plan = ca.GroupedOzaki(handle, descs(probs), moduli=12, split_min_dim=0)
    plan.run()
    torch.cuda.synchronize()
    got = results(probs)
    assert not np.isfinite(got[0][0]).all()
    assert np.isfinite(got[0][1:]).all()
    ref = oracle_result(bad)
    # the non-finite problem is computed by FP64 dot products: the FP64 bound
    assert oracle.rel_fro(got[0][1:], ref[1:]) <= 4 * bad["k"] * 2.0 ** -53
    for p, g in zip(probs[1:], got[1:]):
        assert oracle.rel_fro(g, oracle_result(p)) <= 1e-11
    plan.close()


def test_invalid_plans_are_rejected(handle):
    import torch
    import cosma_amd as ca
    a = torch.zeros((8, 8), dtype=torch.float64, device="cuda")
    good = ("N", "N", 8, 8, 8, 1.0, a.data_ptr(), 8, a.data_ptr(), 8, 0.0, a.data_ptr(), 8)
    with pytest.raises(ca.CosmaAmdError, match="moduli"):
        ca.GroupedOzaki(handle, [good], moduli=21)
    host = np.zeros((8, 8))
    with pytest.raises(ca.CosmaAmdError, match="device memory"):
        ca.GroupedOzaki(handle, [("N", "N", 8, 8, 8, 1.0, host.ctypes.data, 8, a.data_ptr(), 8, 0.0,
                                  a.data_ptr(), 8)])
    empty = ca.GroupedOzaki(handle, [])
    empty.run()
    empty.graph_run()
    empty.close()


def test_small_problems_run_in_fp64_by_default(handle):
    """With the default threshold every problem of dimensions below 2048 runs
    on the FP64 grouped kernel: bit-identical to cosma_amd_dgemm on it, for
    the direct run and for graph replays, with no moduli count chosen."""
    import torch
    import cosma_amd as ca
    probs = make_problems(40, 21, max_dim=300)
    reset_c(probs)
    expected = single_dgemm_results(handle, probs)
    plan = ca.GroupedOzaki(handle, descs(probs))
    info = plan.info()
    assert info["fp64_problems"] == len(probs) and info["split_problems"] == 0, info
    assert info["moduli"] == 0 and info["tiles"] == 0
    plan.tune()
    reset_c(probs)
    plan.run()
    torch.cuda.synchronize()
    assert same(results(probs), expected)
    for _ in range(3):
        reset_c(probs)
        plan.graph_run()
        torch.cuda.synchronize()
        assert same(results(probs), expected)
    assert plan.check()["worst_problem"] == -1
    plan.close()


def test_mixed_routing_at_a_threshold(handle):
    """A threshold inside the batch: problems with min(m, n, k) >= 64 are split
    (bit-identical to cosma_amd_dgemm_ozaki with the plan's count), the others
    run in FP64 (bit-identical to cosma_amd_dgemm), in one run and one graph."""
    import torch
    import cosma_amd as ca
    probs = make_problems(40, 22, max_dim=160)
    threshold, moduli = 64, 14
    split = [min(p["m"], p["n"], p["k"]) >= threshold for p in probs]
    assert 0 < sum(split) < len(probs)
    reset_c(probs)
    exp_oz = single_results(handle, probs, moduli)
    exp_fp = single_dgemm_results(handle, probs)
    expected = [o if s else f for o, f, s in zip(exp_oz, exp_fp, split)]
    plan = ca.GroupedOzaki(handle, descs(probs), moduli=moduli, split_min_dim=threshold)
    info = plan.info()
    assert info["split_problems"] == sum(split) and info["fp64_problems"] == len(probs) - sum(split), info
    assert info["moduli"] == moduli
    reset_c(probs)
    plan.run()
    torch.cuda.synchronize()
    assert same(results(probs), expected)
    for _ in range(2):
        reset_c(probs)
        plan.graph_run()
        torch.cuda.synchronize()
        assert same(results(probs), expected)
    plan.close()

