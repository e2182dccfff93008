"""Pins the CPU oracle (no GPU): golden vectors, the compiled reference,
closed forms, invariants and brute force.  Each test is chosen so that a
plausible mistake in the oracle (a dropped term, a wrong sign or index, a
transposed operand, a read of C when beta == 0) fails it."""
import math
from fractions import Fraction

import numpy as np
import pytest

import oracle
from helpers import (U, gemm_bound, gemm_fixtures, int_fixtures, load_gemm, load_transform,
                     stored_shape, transform_fixtures)


# ------------------------------------------------------------------ RNG
def test_splitmix64_published_vector():
    # SplitMix64 from state 0: the published first outputs of the generator.
    g = 0x9E3779B97F4A7C15
    outs = [oracle.splitmix64((i * g) & (2 ** 64 - 1)) for i in range(3)]
    assert outs == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def test_uniform_is_exact_dyadic_in_range():
    x = oracle.uniform(1000, 3, 5)
    assert x.min() >= -1.0 and x.max() < 1.0
    scaled = (x + 1.0) * 2.0 ** 52        # ((r >> 11) * 2^-53 * 2) * 2^52 = r >> 11, an integer
    assert np.array_equal(scaled, np.floor(scaled))
    assert abs(x.mean()) < 0.05
    # offset = position in the stream
    y = oracle.uniform(1, 3000, 5)
    assert np.array_equal(y.ravel(order="F"), x.ravel(order="F"))


def test_uniform_at_matches_c_generator():
    x = oracle.uniform(50, 40, 77)
    idx = np.arange(50 * 40)
    assert np.array_equal(oracle.uniform_at(77, idx), x.ravel(order="F"))


# --------------------------------------------------------- naive local multiply
@pytest.mark.parametrize("m,n,k", [(4, 4, 4), (3, 4, 5), (17, 9, 31), (1, 1, 1), (64, 3, 7)])
@pytest.mark.parametrize("alpha,beta", [(1.0, 0.0), (1.0, 1.0), (-1.3, 0.5)])
def test_naive_loop_matches_compiled_reference(m, n, k, alpha, beta):
    if oracle.ref_lib() is None:
        pytest.skip("oracle/_ref/liblm_ref.so not built (needs the reference tree)")
    A, B, C = oracle.uniform(m, k, 1), oracle.uniform(k, n, 2), oracle.uniform(m, n, 3)
    c1, c2 = C.copy(order="F"), C.copy(order="F")
    oracle.local_multiply_cpu(A, B, c1, m, n, k, alpha, beta)
    assert oracle.local_multiply_cpu_ref(A, B, c2, m, n, k, alpha, beta)
    assert np.array_equal(c1, c2), "restated loop must be bit-identical to local_multiply.cpp:277-297"


def test_naive_loop_exact_on_integers():
    rng = np.random.default_rng(3)
    m, n, k = 13, 11, 29
    A = np.asfortranarray(rng.integers(-50, 51, (m, k)).astype(np.float64))
    B = np.asfortranarray(rng.integers(-50, 51, (k, n)).astype(np.float64))
    C = np.asfortranarray(rng.integers(-50, 51, (m, n)).astype(np.float64))
    exact = 3 * (A.astype(np.int64) @ B.astype(np.int64)) + 2 * C.astype(np.int64)
    oracle.local_multiply_cpu(A, B, C, m, n, k, 3.0, 2.0)
    assert np.array_equal(C, exact.astype(np.float64))


def test_naive_loop_reads_c_when_beta_zero():
    # Like the reference (Cvalue *= beta), NaN * 0 stays NaN: callers must zero C.
    A, B = oracle.uniform(3, 2, 1), oracle.uniform(2, 3, 2)
    C = np.full((3, 3), np.nan, order="F")
    oracle.local_multiply_cpu(A, B, C, 3, 3, 2, 1.0, 0.0)
    assert np.isnan(C).all()


# ------------------------------------------------------------ BLAS dgemm
@pytest.mark.parametrize("ta", "NT")
@pytest.mark.parametrize("tb", "NT")
def test_dgemm_transposes_are_pure_data_movement(ta, tb):
    # op() must only change where an element is read: dgemm on explicitly
    # transposed storage equals dgemm 'N','N' bit for bit.
    m, n, k = 23, 19, 37
    X, Y, C0 = oracle.uniform(m, k, 4), oracle.uniform(k, n, 5), oracle.uniform(m, n, 6)
    A = X if ta == "N" else np.asfortranarray(X.T)
    B = Y if tb == "N" else np.asfortranarray(Y.T)
    c_nn, c_op = C0.copy(order="F"), C0.copy(order="F")
    oracle.dgemm("N", "N", m, n, k, -0.7, X, m, Y, k, 0.25, c_nn, m)
    oracle.dgemm(ta, tb, m, n, k, -0.7, A, A.shape[0], B, B.shape[0], 0.25, c_op, m)
    assert np.array_equal(c_nn, c_op)


def test_dgemm_agrees_with_naive_loop():
    m, n, k = 31, 27, 45
    A, B, C0 = oracle.uniform(m, k, 7), oracle.uniform(k, n, 8), oracle.uniform(m, n, 9)
    c1, c2 = C0.copy(order="F"), C0.copy(order="F")
    oracle.dgemm("N", "N", m, n, k, 1.5, A, m, B, k, -0.5, c1, m)
    oracle.local_multiply_cpu(A, B, c2, m, n, k, 1.5, -0.5)
    assert oracle.rel_fro(c1, c2) <= 2 * gemm_bound(k)


def test_dgemm_semantics_edge_cases():
    A, B = oracle.uniform(5, 4, 1), oracle.uniform(4, 6, 2)
    C0 = oracle.uniform(5, 6, 3)
    # beta == 0 must not read C (NaN does not propagate)
    C = np.full((5, 6), np.nan, order="F")
    oracle.dgemm("N", "N", 5, 6, 4, 1.0, A, 5, B, 4, 0.0, C, 5)
    assert np.isfinite(C).all()
    # k == 0 and alpha == 0: C <- beta * C
    for k, alpha in ((0, 1.0), (4, 0.0)):
        C = C0.copy(order="F")
        oracle.dgemm("N", "N", 5, 6, k, alpha, A, 5, B, 4, 0.5, C, 5)
        assert np.array_equal(C, 0.5 * C0)
    # m == 0 / n == 0: untouched
    C = C0.copy(order="F")
    oracle.dgemm("N", "N", 0, 6, 4, 1.0, A, 5, B, 4, 0.0, C, 5)
    oracle.dgemm("N", "N", 5, 0, 4, 1.0, A, 5, B, 4, 0.0, C, 5)
    assert np.array_equal(C, C0)
    with pytest.raises(ValueError):
        oracle.dgemm("X", "N", 5, 6, 4, 1.0, A, 5, B, 4, 0.0, C, 5)


def test_dgemm_dd_exact_against_rationals():
    # Tiny case computed in exact rational arithmetic: the double-double
    # result must be the correctly rounded value (within half an ulp... 1 ulp).
    m, n, k = 3, 2, 50
    A, B, C0 = oracle.uniform(m, k, 11), oracle.uniform(k, n, 12), oracle.uniform(m, n, 13)
    alpha, beta = -1.3, 0.7
    C = C0.copy(order="F")
    oracle.dgemm("N", "N", m, n, k, alpha, A, m, B, k, beta, C, m, dd=True)
    for i in range(m):
        for j in range(n):
            s = sum(Fraction(A[i, l]) * Fraction(B[l, j]) for l in range(k))
            exact = Fraction(alpha) * s + Fraction(beta) * Fraction(C0[i, j])
            assert abs(Fraction(C[i, j]) - exact) <= Fraction(math.ulp(float(exact)))


# ---------------------------------------------------------- golden fixtures
@pytest.mark.parametrize("name", gemm_fixtures())
def test_oracle_against_golden_host_blas(name):
    d = load_gemm(name)
    m, n, k = d["m"], d["n"], d["k"]
    for idx, (alpha, beta) in enumerate(d["alphabeta"]):
        C = d["C0"].copy(order="F")
        oracle.dgemm(d["transa"], d["transb"], m, n, k, alpha, d["A"], d["A"].shape[0], d["B"], d["B"].shape[0],
                     beta, C, m)
        assert oracle.rel_fro(C, d[f"C_blas_{idx}"]) <= gemm_bound(k), (name, idx)
        if f"C_dd_{idx}" in d:
            # host BLAS and the oracle are both within the FP64 bound of the dd value
            assert oracle.rel_fro(d[f"C_blas_{idx}"], d[f"C_dd_{idx}"]) <= gemm_bound(k)
            assert oracle.rel_fro(C, d[f"C_dd_{idx}"]) <= gemm_bound(k)


@pytest.mark.parametrize("name", int_fixtures())
def test_oracle_exact_on_integer_golden(name):
    d = load_gemm(name)
    m, n, k = d["m"], d["n"], d["k"]
    for dd in (False, True):
        C = np.zeros((m, n), order="F")
        oracle.dgemm(d["transa"], d["transb"], m, n, k, 1.0, d["A"], d["A"].shape[0], d["B"], d["B"].shape[0],
                     0.0, C, m, dd=dd)
        assert np.array_equal(C, d["C_exact"])


# ------------------------------------------------------------ block transform
def test_block_transform_index_map_closed_form():
    rows, cols, lds, ldd = 7, 5, 9, 8
    src = np.asfortranarray(np.arange(lds * rows, dtype=np.float64).reshape((lds, rows), order="F"))
    dst = np.full((ldd, cols), -1.0, order="F")
    oracle.block_transform("T", rows, cols, 1.0, src, lds, 0.0, dst, ldd)
    for i in range(rows):
        for j in range(cols):
            assert dst[i, j] == src[j, i] == j + i * lds
    assert (dst[rows:, :] == -1.0).all(), "padding rows of dst must be untouched"


def test_block_transform_alpha_beta_and_nan():
    src, dst0 = oracle.uniform(6, 4, 1), oracle.uniform(6, 4, 2)
    d = dst0.copy(order="F")
    oracle.block_transform("N", 6, 4, -1.3, src, 6, 0.5, d, 6)
    assert np.allclose(d, -1.3 * src + 0.5 * dst0, rtol=2 * U, atol=0)
    d = np.full((6, 4), np.nan, order="F")
    oracle.block_transform("N", 6, 4, 2.0, src, 6, 0.0, d, 6)
    assert np.array_equal(d, 2.0 * src)


@pytest.mark.parametrize("name", transform_fixtures())
def test_block_transform_golden(name):
    d = load_transform(name)
    out = d["dst0"].copy(order="F")
    oracle.block_transform(d["op"], d["rows"], d["cols"], d["alpha"], d["src"], d["lds"], d["beta"], out, d["ldd"])
    assert np.array_equal(out, d["dst"])
    if d["alpha"] == 1.0 and d["beta"] == 0.0:
        view = d["src"][:d["rows"], :d["cols"]] if d["op"] == "N" else d["src"][:d["cols"], :d["rows"]].T
        assert np.array_equal(out[:d["rows"], :d["cols"]], view)


# ----------------------------------------------------------- ScaLAPACK maps
@pytest.mark.parametrize("n,nb,P,src", [(100, 7, 3, 0), (128, 32, 4, 1), (1000, 128, 1, 0), (5, 8, 2, 1),
                                         (16384, 128, 1, 0), (83, 10, 4, 3)])
def test_numroc_matches_brute_force(n, nb, P, src):
    counts = [0] * P
    for g in range(n):
        p, _ = oracle.global_to_local(g, nb, src, P)
        counts[p] += 1
    assert [oracle.numroc(n, nb, p, src, P) for p in range(P)] == counts
    if P == 1:
        assert oracle.numroc(n, nb, 0, 0, 1) == n


def test_global_to_local_is_a_bijection():
    n, nb, P, src = 200, 9, 3, 2
    seen = {p: set() for p in range(P)}
    for g in range(n):
        p, l = oracle.global_to_local(g, nb, src, P)
        assert l not in seen[p]
        seen[p].add(l)
    for p in range(P):
        assert seen[p] == set(range(oracle.numroc(n, nb, p, src, P)))


def test_block_cyclic_gather_one_process_is_identity():
    m, n, lld = 40, 30, 45
    local = oracle.uniform(lld, n + 5, 3)
    dst = np.zeros((m, n), order="F")
    oracle.block_cyclic_gather(local, lld, 8, 8, 0, 0, 1, 1, 0, 0, 2, 3, m - 2, n - 3, dst, m)
    assert np.array_equal(dst[:m - 2, :n - 3], local[2:m, 3:n])


def test_block_cyclic_gather_2x2_grid_reassembles():
    M, N, mb, nb, P, Q = 21, 17, 4, 3, 2, 2
    G = oracle.uniform(M, N, 9)
    out = np.zeros((M, N), order="F")
    for pr in range(P):
        for pc in range(Q):
            lr, lc = oracle.numroc(M, mb, pr, 0, P), oracle.numroc(N, nb, pc, 0, Q)
            local = np.zeros((max(lr, 1), max(lc, 1)), order="F")
            for g in range(M):
                p, l = oracle.global_to_local(g, mb, 0, P)
                if p != pr:
                    continue
                for h in range(N):
                    q, c = oracle.global_to_local(h, nb, 0, Q)
                    if q == pc:
                        local[l, c] = G[g, h]
            oracle.block_cyclic_gather(local, local.shape[0], mb, nb, 0, 0, P, Q, pr, pc, 0, 0, M, N, out, M)
    assert np.array_equal(out, G)


def test_host_blas_correct_with_torch_loaded():
    # The CPU baseline runs in the same process as torch; MKL must not clash
    # with torch's OpenMP runtime (it returned wrong values when it did).
    import torch  # noqa: F401
    m, n, k = 700, 650, 900
    A, B = oracle.uniform(m, k, 1), oracle.uniform(k, n, 2)
    C = np.zeros((m, n), order="F")
    oracle.hostblas.set_threads(4)
    oracle.hostblas.dgemm("N", "N", m, n, k, 1.0, A, m, B, k, 0.0, C, m)
    ref = np.zeros((m, n), order="F")
    oracle.dgemm("N", "N", m, n, k, 1.0, A, m, B, k, 0.0, ref, m)
    assert oracle.rel_fro(C, ref) <= gemm_bound(k)


@pytest.mark.parametrize("ta,tb", [("N", "N"), ("T", "N"), ("N", "T"), ("T", "T")])
def test_sampled_lines_helper_matches_full_product(ta, tb):
    """The full-size checks take whole rows / columns of op(A) op(B) from the
    oracle through sub-block addresses; they must equal the same lines of the
    oracle's full product."""
    from oracle.sampled import dd_lines, lines_error
    m, n, k = 37, 29, 45
    A = oracle.uniform(*((m, k) if ta == "N" else (k, m)), 71)
    B = oracle.uniform(*((k, n) if tb == "N" else (n, k)), 72)
    full = np.zeros((m, n), order="F")
    oracle.dgemm(ta, tb, m, n, k, 1.0, A, A.shape[0], B, B.shape[0], 0.0, full, m, dd=True)
    rows, cols = (0, 17, 36), (0, 28)
    r, c = dd_lines(ta, tb, m, n, k, A.ctypes.data, A.shape[0], B.ctypes.data, B.shape[0], rows, cols, threads=3)
    for i in rows:
        assert np.array_equal(r[i], full[i, :])
    for j in cols:
        assert np.array_equal(c[j], full[:, j])
    assert lines_error(r, c, {i: full[i, :] for i in rows}, {j: full[:, j] for j in cols}) == 0.0
