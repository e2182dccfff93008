"""Split-precision (Ozaki) DGEMM on the INT8 matrix cores, through the C ABI.

Pinned three ways:
  * exactness: integer-valued inputs that the power-of-two scaling keeps
    exact give the exact product, bit for bit, for both methods;
  * emulation: the moduli method must return the exact product of the
    P-bit-rounded inputs (computed here with Python integers) to 2 ulp, and
    the slices method must match a float64 emulation of its digit splitting
    and recombination bit for bit;
  * accuracy: against the double-double oracle, the default tolerance gives
    relative Frobenius error <= 1e-14 for uniform inputs.
"""
import math

import numpy as np
import pytest

import oracle
from helpers import from_dev, gemm_bound, stored_shape, to_dev

pytestmark = pytest.mark.gpu

MODULI = [256, 255, 253, 251, 247, 241, 239, 233, 229, 227, 223, 217, 211, 199, 197, 193, 191, 181, 179, 173]


def op(X, t):
    return X if t == "N" else np.asfortranarray(X.T)


def run_ozaki(handle, ta, tb, m, n, k, alpha, A, B, beta, C0, method, tol=1e-14, count=0, nan_c=False):
    import torch
    dA, lda = to_dev(A)
    dB, ldb = to_dev(B)
    dC, ldc = to_dev(C0)
    if nan_c:
        dC.fill_(float("nan"))
    info = handle.dgemm_ozaki(ta, tb, m, n, k, alpha, dA.data_ptr(), lda, dB.data_ptr(), ldb, beta,
                              dC.data_ptr(), ldc, method=method, tolerance=tol, count=count)
    torch.cuda.synchronize()
    return from_dev(dC, m), info


def row_exponents(X):
    """frexp exponent of the max |entry| of each row (0 for zero rows)."""
    mx = np.abs(X).max(axis=1)
    return np.where(mx > 0, np.frexp(mx)[1], 0)


def emulate_moduli(Aop, Bop, P, alpha):
    E = row_exponents(Aop)
    F = row_exponents(Bop.T)
    Ai = np.rint(np.ldexp(Aop, (P - E)[:, None])).astype(np.int64)
    Bi = np.rint(np.ldexp(Bop, (P - F)[None, :])).astype(np.int64)
    Cexact = Ai.astype(object) @ Bi.astype(object)          # Python integers: exact
    m, n = Cexact.shape
    out = np.empty((m, n))
    for i in range(m):
        for j in range(n):
            out[i, j] = alpha * math.ldexp(float(Cexact[i, j]), int(E[i] + F[j] - 2 * P))
    return out


def emulate_slices(Aop, Bop, s, alpha):
    def exps(X):
        mx = np.abs(X).max(axis=1)
        f, e = np.frexp(mx)
        e = np.where(f * 128.0 >= 127.5, e + 1, e)
        return np.where(mx > 0, e, 0)

    def digits(X, e):
        r = np.ldexp(X, -e[:, None])
        out = []
        for _ in range(s):
            y = r * 128.0
            d = np.rint(y)
            r = y - d
            out.append(d.astype(np.int64))
        return out

    E, F = exps(Aop), exps(Bop.T)
    da, db = digits(Aop, E), digits(Bop.T, F)
    acc = None
    for d in range(2, s + 2):
        tot = 0
        for t in range(1, s + 1):
            u = d - t
            if 1 <= u <= s:
                tot = tot + da[t - 1] @ db[u - 1].T
        val = np.ldexp(tot.astype(np.float64), -7 * d)
        acc = val if acc is None else acc + val
    return alpha * np.ldexp(acc, E[:, None] + F[None, :])


@pytest.mark.parametrize("method,count", [(0, 0), (1, 3), (2, 0)])
@pytest.mark.parametrize("ta,tb", [("N", "N"), ("T", "N"), ("N", "T")])
def test_integer_inputs_are_exact(handle, method, count, ta, tb):
    m, n, k = 130, 70, 300
    rng = np.random.default_rng(1)
    X = rng.integers(-1024, 1025, (m, k)).astype(np.float64)
    Y = rng.integers(-1024, 1025, (k, n)).astype(np.float64)
    A, B = np.asfortranarray(op(X, ta)), np.asfortranarray(op(Y, tb))
    got, info = run_ozaki(handle, ta, tb, m, n, k, 1.0, A, B, 0.0, np.zeros((m, n), order="F"), method,
                          count=count, nan_c=True)
    exact = (X.astype(np.int64) @ Y.astype(np.int64)).astype(np.float64)
    assert info["method"] == method
    assert np.array_equal(got, exact)


@pytest.mark.parametrize("count", [2, 5, 9, 14, 16, 18, 20])
@pytest.mark.parametrize("ta,tb", [("N", "N"), ("T", "T")])
def test_moduli_equals_exact_product_of_rounded_inputs(handle, count, ta, tb):
    m, n, k = 37, 45, 129
    X, Y = oracle.uniform(m, k, count), oracle.uniform(k, n, 100 + count)
    X[3, :] *= 2.0 ** 40        # a row with a large exponent
    Y[:, 7] *= 2.0 ** -30       # a column with a small one
    X[5, :] = 0.0               # a zero row
    A, B = np.asfortranarray(op(X, ta)), np.asfortranarray(op(Y, tb))
    got, info = run_ozaki(handle, ta, tb, m, n, k, -1.3, A, B, 0.0, np.zeros((m, n), order="F"), 0, count=count)
    assert info["count"] == count
    ref = emulate_moduli(X, Y, info["bits"], -1.3)
    diff = np.abs(got - ref)
    assert (diff <= 2 * np.spacing(np.abs(ref))).all(), diff.max()


@pytest.mark.parametrize("method", [0, 2])
def test_norm_bound_raises_bits_and_stays_exact(handle, method):
    """Rows of op(A) / columns of op(B) with one entry at the maximum and the
    rest far below it: their 2-norms bound |A'B'| much more tightly than
    k 2^(2P), so the bits P chosen for 8 moduli exceed the entrywise limit
    floor((log2 M - 1 - log2 k) / 2), and the result must still be the exact
    product of the P-bit-rounded inputs (|A'B'| < M / 2 is what makes the
    residues recombine exactly)."""
    m, n, k, count = 48, 40, 512, 8
    rng = np.random.default_rng(3)
    X = rng.uniform(-1, 1, (m, k)) * 2.0 ** -12
    Y = rng.uniform(-1, 1, (k, n)) * 2.0 ** -12
    X[np.arange(m), rng.integers(0, k, m)] = 1.0
    Y[rng.integers(0, k, n), np.arange(n)] = -1.0
    X, Y = np.asfortranarray(X), np.asfortranarray(Y)
    got, info = run_ozaki(handle, "N", "N", m, n, k, 1.0, X, Y, 0.0, np.zeros((m, n), order="F"), method,
                          count=count)
    log2M = sum(math.log2(x) for x in MODULI[:count])
    entrywise = math.floor((log2M - 1 - math.log2(k) - 1e-6) / 2)
    assert info["bits"] > entrywise, (info, entrywise)
    ref = emulate_moduli(X, Y, info["bits"], 1.0)
    assert (np.abs(got - ref) <= 2 * np.spacing(np.abs(ref))).all()


@pytest.mark.parametrize("s", [1, 2, 4, 7, 8])
def test_slices_match_emulation_bit_for_bit(handle, s):
    m, n, k = 41, 33, 200
    X, Y = oracle.uniform(m, k, 7 + s), oracle.uniform(k, n, 70 + s)
    X[2, :] *= 2.0 ** 20
    X[9, :] = 0.0
    got, info = run_ozaki(handle, "N", "N", m, n, k, 0.75, X, Y, 0.0, np.zeros((m, n), order="F"), 1, count=s)
    assert info["count"] == s and info["int8_products"] == s * (s + 1) // 2
    assert np.array_equal(got, emulate_slices(X, Y, s, 0.75))


@pytest.mark.parametrize("method", [0, 1, 2])
@pytest.mark.parametrize("m,n,k,ta,tb", [(256, 256, 1024, "N", "N"), (300, 200, 777, "T", "N"),
                                          (129, 257, 2048, "N", "T"), (512, 64, 4096, "T", "T")])
def test_default_tolerance_against_double_double(handle, method, m, n, k, ta, tb):
    A = oracle.uniform(*stored_shape(ta, m, k), 1)
    B = oracle.uniform(*stored_shape(tb, k, n), 2)
    C0 = oracle.uniform(m, n, 3)
    got, info = run_ozaki(handle, ta, tb, m, n, k, 1.0, A, B, 0.5, C0, method)
    ref = C0.copy(order="F")
    oracle.dgemm(ta, tb, m, n, k, 1.0, A, A.shape[0], B, B.shape[0], 0.5, ref, m, dd=True)
    err = oracle.rel_fro(got, ref)
    assert err <= 1e-14, (err, info)
    assert err <= info["error_bound"], (err, info)
    assert info["tolerance_met"] == 1 and info["passes"] == 1, info


@pytest.mark.parametrize("count", [2, 9, 14, 20])
@pytest.mark.parametrize("m,n,k,ta,tb", [(37, 45, 129, "N", "N"), (300, 260, 1800, "T", "N"),
                                          (130, 70, 4000, "N", "T")])
def test_bf16_residues_bit_identical_to_int8(handle, count, m, n, k, ta, tb):
    """The BF16 variant holds the same residues and its FP32 sums are exact
    (reduced modulo m every 896 k), so it must equal the INT8 path bit for
    bit; k = 1800 and 4000 cross two and four reduction points."""
    import cosma_amd as ca
    A = oracle.uniform(*stored_shape(ta, m, k), 11 + count)
    B = oracle.uniform(*stored_shape(tb, k, n), 12 + count)
    A[(0, 3), (1, 2)] *= 2.0 ** 30       # entries far above the rest of their row / column
    C0 = oracle.uniform(m, n, 13)
    got8, info8 = run_ozaki(handle, ta, tb, m, n, k, -0.7, A, B, 0.5, C0, ca.OZAKI_MODULI, count=count)
    got16, info16 = run_ozaki(handle, ta, tb, m, n, k, -0.7, A, B, 0.5, C0, ca.OZAKI_MODULI_BF16, count=count)
    assert info16["method"] == ca.OZAKI_MODULI_BF16 and info16["count"] == count
    assert info16["bits"] == info8["bits"] and info16["int8_products"] == count
    assert np.array_equal(got16, got8)


def test_bf16_long_sums_of_large_odd_products_stay_exact(handle):
    """Constant operands: every residue plane is constant, so each FP32 sum
    adds k copies of the same product r_A * r_B.  For the chosen values some
    of these products are odd and k * |r_A r_B| exceeds 2^24, where FP32
    addition without the periodic reduction would drop bits (checked below
    on the CPU); the result must still be the exact product of the rounded
    inputs."""
    import cosma_amd as ca
    m, n, k, count = 64, 48, 4096, 14
    x, y = 0.8137, 0.6071
    A = np.full((m, k), x, order="F")
    B = np.full((k, n), y, order="F")
    got, info = run_ozaki(handle, "N", "N", m, n, k, 1.0, A, B, 0.0, np.zeros((m, n), order="F"),
                          ca.OZAKI_MODULI_BF16, count=count)
    P = info["bits"]
    ai, bi = int(np.rint(np.ldexp(x, P))), int(np.rint(np.ldexp(y, P)))

    def sym(v, md):
        r = v % md
        return r - md if r > (md - 1) // 2 else r
    risky = [md for md in MODULI[:count]
             if (sym(ai, md) * sym(bi, md)) % 2 and k * abs(sym(ai, md) * sym(bi, md)) > 2 ** 24]
    assert risky, "the operands do not exercise sums beyond 2^24"
    md = risky[0]
    prod = np.float32(sym(ai, md) * sym(bi, md))
    naive = np.float32(0)
    for _ in range(k):
        naive = np.float32(naive + prod)
    assert int(naive) != k * int(prod)          # FP32 alone would be inexact here
    exact = math.ldexp(float(k * ai * bi), -2 * P)
    assert np.all(np.abs(got - exact) <= 2 * np.spacing(exact)), (got[0, 0], exact)
    got8, _ = run_ozaki(handle, "N", "N", m, n, k, 1.0, A, B, 0.0, np.zeros((m, n), order="F"),
                        ca.OZAKI_MODULI, count=count)
    assert np.array_equal(got, got8)


def test_count_from_tolerance_is_monotone(handle):
    m = n = 64
    k = 256
    A, B = oracle.uniform(m, k, 1), oracle.uniform(k, n, 2)
    counts, errs = [], []
    for tol in (1e-4, 1e-8, 1e-12, 1e-15):
        got, info = run_ozaki(handle, "N", "N", m, n, k, 1.0, A, B, 0.0, np.zeros((m, n), order="F"), 0, tol=tol)
        ref = np.zeros((m, n), order="F")
        oracle.dgemm("N", "N", m, n, k, 1.0, A, m, B, k, 0.0, ref, m, dd=True)
        counts.append(info["count"])
        errs.append(oracle.rel_fro(got, ref))
        assert errs[-1] <= tol, (tol, errs[-1], info)
    assert counts == sorted(counts)


def test_lognormal_inputs_error_reported(handle):
    """phi = 4 inputs: entries spread over ~2^+-17 within rows; the error is
    reported (not bounded by 1e-14 for few moduli), and more moduli reduce it."""
    import torch
    import cosma_amd as ca
    m = n = 128
    k = 512
    A = torch.empty((k, m), dtype=torch.float64, device="cuda")
    B = torch.empty((n, k), dtype=torch.float64, device="cuda")
    ca.fill_lognormal(A.data_ptr(), m, k, m, 5, 4.0)
    ca.fill_lognormal(B.data_ptr(), k, n, k, 6, 4.0)
    torch.cuda.synchronize()
    Ah, Bh = from_dev(A, m), from_dev(B, k)
    ref = np.zeros((m, n), order="F")
    oracle.dgemm("N", "N", m, n, k, 1.0, Ah, m, Bh, k, 0.0, ref, m, dd=True)
    errs = []
    for count in (8, 12, 16):
        got, info = run_ozaki(handle, "N", "N", m, n, k, 1.0, Ah, Bh, 0.0, np.zeros((m, n), order="F"), 0,
                              count=count)
        errs.append(oracle.rel_fro(got, ref))
    assert errs[0] > errs[1] > errs[2]
    assert errs[2] < 1e-13


def test_default_tolerance_holds_for_spread_exponents(handle):
    """phi = 4 inputs (entries spread over ~2^+-17 below their row maxima): the
    count chosen from the data must still meet the default 1e-14, and it must
    be larger than for uniform inputs of the same shape."""
    import torch
    import cosma_amd as ca
    m = n = 192
    k = 768
    A = torch.empty((k, m), dtype=torch.float64, device="cuda")
    B = torch.empty((n, k), dtype=torch.float64, device="cuda")
    ca.fill_lognormal(A.data_ptr(), m, k, m, 7, 4.0)
    ca.fill_lognormal(B.data_ptr(), k, n, k, 8, 4.0)
    torch.cuda.synchronize()
    Ah, Bh = from_dev(A, m), from_dev(B, k)
    ref = np.zeros((m, n), order="F")
    oracle.dgemm("N", "N", m, n, k, 1.0, Ah, m, Bh, k, 0.0, ref, m, dd=True)
    got, info = run_ozaki(handle, "N", "N", m, n, k, 1.0, Ah, Bh, 0.0, np.zeros((m, n), order="F"), 0)
    err = oracle.rel_fro(got, ref)
    assert err <= 1e-14, (err, info)
    assert err <= info["error_bound"], (err, info)
    U = oracle.uniform(m, k, 1)
    _, info_u = run_ozaki(handle, "N", "N", m, n, k, 1.0, U, oracle.uniform(k, n, 2), 0.0,
                          np.zeros((m, n), order="F"), 0)
    assert info["count"] > info_u["count"], (info, info_u)


@pytest.mark.parametrize("method", [0, 2])
def test_nonfinite_inputs_fall_back_to_fp64(handle, method):
    m, n, k = 40, 30, 50
    A, B = oracle.uniform(m, k, 1), oracle.uniform(k, n, 2)
    A[3, 4] = np.inf
    got, info = run_ozaki(handle, "N", "N", m, n, k, 1.0, A, B, 0.0, np.zeros((m, n), order="F"), method)
    assert info["method"] == -1
    assert np.isinf(got[3]).any() or np.isnan(got[3]).any()
    assert np.isfinite(np.delete(got, 3, axis=0)).all()


@pytest.mark.parametrize("method", [0, 1, 2])
def test_semantics_beta_and_degenerate(handle, method):
    m, n, k = 50, 40, 64
    A, B, C0 = oracle.uniform(m, k, 1), oracle.uniform(k, n, 2), oracle.uniform(m, n, 3)
    got, _ = run_ozaki(handle, "N", "N", m, n, k, 1.0, A, B, 0.0, C0, method, nan_c=True)
    assert np.isfinite(got).all()
    got, _ = run_ozaki(handle, "N", "N", m, n, 0, 1.0, A, B, 0.25, C0, method)
    assert np.array_equal(got, 0.25 * C0)
    got, _ = run_ozaki(handle, "N", "N", m, n, k, 0.0, A, B, 2.0, C0, method)
    assert np.array_equal(got, 2.0 * C0)


def test_invalid_arguments(handle):
    import cosma_amd as ca
    with pytest.raises(ca.CosmaAmdError, match="method"):
        handle.dgemm_ozaki("N", "N", 4, 4, 4, 1.0, 16, 4, 16, 4, 0.0, 16, 4, method=5)
    with pytest.raises(ca.CosmaAmdError, match="moduli"):
        handle.dgemm_ozaki("N", "N", 4, 4, 4, 1.0, 16, 4, 16, 4, 0.0, 16, 4, method=0, count=21)
