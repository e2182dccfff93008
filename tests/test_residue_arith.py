"""The INT8 product kernel's residue arithmetic (i8gemm.hpp: sym_mod_odd_bits,
sym_mod_odd_bits_wide), restated in numpy float32 and checked against exact
integer residues.

Every float32 step of the two functions is either exact (conversions of
integers below 2^24, h * c16 + l below 2^22, q * m, the sums with 1.5 * 2^23)
or one correctly rounded product (y * fl(1/m)) followed by round-half-even,
so numpy reproduces the device's results bit for bit.  The check covers all
20 moduli of the split-precision DGEMM (256 and the odd ones), every x in
[-2^20, 2^20] and random x up to the bounds each form is used for (|x| < 2^29
while K <= 32768 bytes; any int32 for the wide form).  The GPU tests check
the kernel itself: its residues feed exact products that are compared with
exact integer results."""
import numpy as np

MODULI = [256, 255, 253, 251, 247, 241, 239, 233, 229, 227, 223, 217, 211, 199, 197, 193, 191, 181, 179, 173]
MAGIC = np.float32(12582912.0)   # 1.5 * 2^23


def fast_bits(x, m):
    x = x.astype(np.int32)
    c16, inv, neg = np.float32(65536 % m), np.float32(1.0) / np.float32(m), np.float32(-m)
    yf = (x >> 16).astype(np.float32) * c16 + (x & 0xFFFF).astype(np.float32)
    q = np.rint(yf * inv)
    t = (yf + MAGIC) + q * neg
    return (t.view(np.uint32) & 0xFF).astype(np.uint8).view(np.int8)


def wide_bits(x, m):
    x = x.astype(np.int32)
    inv, neg = np.float32(1.0) / np.float32(m), np.float32(-m)
    qm = np.rint((x >> 16).astype(np.float32) * inv) * neg
    x2 = (x.astype(np.int64) + (qm.astype(np.int64) << 16)).astype(np.int32)
    return fast_bits(x2, m)


def symmetric(x, m):
    r = np.mod(x.astype(np.int64), m)
    r = np.where(r > (m - 1) // 2, r - m, r)
    return r.astype(np.int8)


def test_fast_form_all_moduli():
    rng = np.random.default_rng(5)
    dense = np.arange(-(1 << 20), (1 << 20) + 1, dtype=np.int64)
    sample = rng.integers(-(1 << 29) + 1, 1 << 29, 1 << 20)
    edges = np.array([(1 << 29) - 1, -(1 << 29) + 1, 65535, 65536, -65536, -65537, 0, 1, -1])
    x = np.concatenate([dense, sample, edges])
    for m in MODULI:
        assert np.array_equal(fast_bits(x, m), symmetric(x, m)), m


def test_wide_form_all_moduli():
    rng = np.random.default_rng(6)
    sample = rng.integers(-(1 << 31), 1 << 31, 1 << 20)
    edges = np.array([(1 << 31) - 1, -(1 << 31), (1 << 29), -(1 << 29), 65535, -65537, 0])
    x = np.concatenate([np.arange(-(1 << 18), 1 << 18), sample, edges])
    for m in MODULI:
        assert np.array_equal(wide_bits(x, m), symmetric(x, m)), m


def test_a_wrong_quotient_would_be_caught():
    # the fast form applied beyond its range (|x| near 2^31) must fail for some
    # odd modulus: the check above is sensitive to a quotient off by one
    rng = np.random.default_rng(7)
    x = rng.integers(-(1 << 31), 1 << 31, 1 << 18)
    assert any(not np.array_equal(fast_bits(x, m), symmetric(x, m)) for m in MODULI[1:])
