// Shared pieces of the split-precision (moduli method) DGEMM: the moduli, the
// CRT constants, and the device steps used by both the single-problem path
// (ozaki.hip) and the grouped one (grouped_ozaki.hip).
#pragma once
#include "i8gemm.hpp"

namespace cosma_amd {

// pairwise-coprime moduli, largest first (256 = 2^8; 255 = 3*5*17; 253 = 11*23;
// 247 = 13*19; 217 = 7*31; the rest are primes)
inline constexpr int kModuli[20] = {256, 255, 253, 251, 247, 241, 239, 233, 229, 227,
                                    223, 217, 211, 199, 197, 193, 191, 181, 179, 173};
constexpr int MAX_MODULI = 20;
// CRT limbs: 4 limbs of 41 bits cover M < 2^164 (20 moduli: M ~ 2^155.4), and a
// limb sum of 20 terms |c_l| * w < 20 * 2^7 * 2^41 < 2^53 stays exact in FP64.
constexpr int LIMB = 41;
constexpr int NLIMB = 4;
// Largest P (bits of the integer inputs) for which the residues stay exact:
// a' / m_l < 2^51 for every modulus (>= 173) keeps the split's FP64 quotient
// within one of the true one, so a single correction step is exact.
constexpr int MAX_P = 58;
// Exponent given to an all-zero row or column: its entries and products are
// zero whatever the scaling, and it drops out of the error estimate.
constexpr int ZERO_EXPONENT = -30000;

struct CrtConst {
    int n;
    int m[MAX_MODULI];
    double inv_m[MAX_MODULI];
    double w[MAX_MODULI][NLIMB];   // CRT weights W_l in 41-bit limbs (exact in FP64)
    int64_t mlimb[NLIMB];          // M in 41-bit limbs
    double inv_M;                  // 1 / M
};

double log2_moduli(int n);
// Largest P with k * 2^(2P) < M / 2 for the first n moduli.
int moduli_bits(int n, int64_t k);
// M / 2 for the first n moduli, reduced by the same relative margin as
// moduli_bits: the bound the integer product must stay below.
double moduli_half_limit(int n);

// Largest P in (pmin, MAX_P] with (2^P sa + h)(2^P sb + h) < limit, else
// pmin.  With sa >= ||a_i|| 2^-E_i for every row of op(A), sb likewise for
// the columns of op(B), h = sqrt(k) / 2 and limit = moduli_half_limit(n),
// the integer inputs a' = rint(a 2^(P - E)) satisfy ||a'_i|| <= 2^P sa + h
// (each entry moves by at most 1/2), so by Cauchy-Schwarz |A'B'| stays
// inside (-M/2, M/2).  Plain IEEE arithmetic only, so the host and the
// device agree on the result.
__host__ __device__ inline int norm_bits(double limit, double h, double sa, double sb, int pmin) {
    for (int P = MAX_P; P > pmin; --P) {
        const double lhs = (ldexp(sa, P) + h) * (ldexp(sb, P) + h);
        if (lhs < limit) return P;   // false for NaN / inf
    }
    return pmin;
}

// Exponent range in which the squares of a vector's largest entries neither
// overflow nor underflow in FP64, so its sum of squares bounds its norm.
constexpr int NORM_EXP_LIMIT = 480;
void crt_constants(int n, CrtConst &cc);
// CRT limbs needed for n moduli: 3 while M < 2^122, else 4.
inline int crt_limbs(int n) { return log2_moduli(n) < 3 * LIMB - 1 ? 3 : NLIMB; }

// Row exponent from the max: max < 2^E.  For slices the leading digit must
// stay within [-127, 127]: rows whose scaled max would round to 128 take E + 1.
__device__ __forceinline__ int row_exponent(unsigned long long bits, int slices) {
    const double mx = __longlong_as_double((long long)bits);
    if (mx == 0.0) return ZERO_EXPONENT;
    int e;
    const double f = frexp(mx, &e);   // mx = f * 2^e, f in [0.5, 1)
    if (slices && f * 128.0 >= 127.5) ++e;
    return e;
}

// 16 residues to a plane: 16 int8 bytes, or (BF16) 16 BF16 numbers of the
// same integer values in 32 bytes (integers of at most 8 significant bits are
// exact in BF16: the top half of their FP32 pattern).
template <bool BF16>
__device__ __forceinline__ void store_residues16(const i32x4 &q, int8_t *dst) {
    if constexpr (BF16) {
        union { i32x4 q; int8_t b[16]; } u;
        u.q = q;
        uint32_t w[8];
#pragma unroll
        for (int j = 0; j < 8; ++j)
            w[j] = (__float_as_uint((float)u.b[2 * j]) >> 16) | (__float_as_uint((float)u.b[2 * j + 1]) & 0xffff0000u);
        *reinterpret_cast<i32x4 *>(dst) = i32x4{(int)w[0], (int)w[1], (int)w[2], (int)w[3]};
        *reinterpret_cast<i32x4 *>(dst + 16) = i32x4{(int)w[4], (int)w[5], (int)w[6], (int)w[7]};
    } else {
        *reinterpret_cast<i32x4 *>(dst) = q;
    }
}

// Symmetric residues of 16 consecutive entries v (scaled by 2^(P - e) and
// rounded to integers) modulo the first `count` moduli: 16 bytes (BF16: 32)
// per plane at dst + l * plane.
template <bool BF16 = false>
__device__ __forceinline__ void moduli_residues16(const double (&v)[16], int e, int P, int count, const int *mods,
                                                  const double *inv_mods, int8_t *dst, int64_t plane) {
    double iv[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) iv[j] = rint(ldexp(v[j], P - e));
    if (P <= 50) {
        // |iv| <= 2^50: iv * (1/m) is within |iv| / m * 2^-52 < 1 / (2m) of
        // iv / m, closer than iv / m ever lies to a half-integer it does
        // not sit on, so q = rint(iv / m) exactly and r = iv - m q lies in
        // [-m/2, m/2]; only m = 256 reaches +-128, and 128 wraps to -128,
        // the same residue.
        for (int l = 0; l < count; ++l) {
            const double m = (double)mods[l], im = inv_mods[l];
            union { i32x4 q; int8_t b[16]; } u;
#pragma unroll
            for (int j = 0; j < 16; ++j) u.b[j] = (int8_t)(int)fma(-m, rint(iv[j] * im), iv[j]);
            store_residues16<BF16>(u.q, dst + (int64_t)l * plane);
        }
    } else {
        for (int l = 0; l < count; ++l) {
            const double m = (double)mods[l], im = inv_mods[l];
            const int hi = (mods[l] - 1) / 2, lo = -(mods[l] / 2);
            union { i32x4 q; int8_t b[16]; } u;
#pragma unroll
            for (int j = 0; j < 16; ++j) {
                int r = (int)fma(-m, rint(iv[j] * im), iv[j]);   // exact: FMA of integers
                if (r > hi) r -= mods[l];
                if (r < lo) r += mods[l];
                u.b[j] = (int8_t)r;
            }
            store_residues16<BF16>(u.q, dst + (int64_t)l * plane);
        }
    }
}

// Exact value of the integer whose residues produced the limb sums s (S =
// sum_x s_x 2^(41 x) = sum_l c_l W_l), reduced into (-M/2, M/2), as the
// nearest FP64.  All arithmetic is exact FP64 on integers below 2^53:
//   limb sums    |s| <= 20 * 128 * 2^41 < 2^49.4,
//   quotient     |q| <= 20 * 128 (S / M with |c_l| <= 128, W_l < M),
//   q * limb     < 2^11.4 * 2^41 = 2^52.4, so t = s - q * limb is one exact
//                FMA, |t| < 2^52.5, and carries below 2^11.5 keep the next
//                limb below 2^53.
// q = round(S / M): |A'B'| < M / 2 by the choice of P, and sd * inv_M is
// within ~2^-40 of S / M, so the rounding cannot fall on the wrong side of a
// tie.
template <int NL>
__device__ __forceinline__ double crt_value(const double (&s)[NL], const CrtConst &cc) {
    constexpr double TWO41 = 2199023255552.0, INV_TWO41 = 1.0 / 2199023255552.0;
    double scale[NL];
#pragma unroll
    for (int x = 0; x < NL; ++x) scale[x] = x == 0 ? 1.0 : scale[x - 1] * TWO41;
    double sd = s[NL - 1] * scale[NL - 1];
#pragma unroll
    for (int x = NL - 2; x >= 0; --x) sd += s[x] * scale[x];
    const double q = rint(sd * cc.inv_M);
    double t[NL];
#pragma unroll
    for (int x = 0; x < NL; ++x) t[x] = fma(-q, (double)cc.mlimb[x], s[x]);
#pragma unroll
    for (int x = 0; x < NL - 1; ++x) {    // normalise: limbs below the top in [0, 2^41)
        const double carry = floor(t[x] * INV_TWO41);
        t[x] = fma(-carry, TWO41, t[x]);
        t[x + 1] += carry;
    }
    // value = sum_x t_x 2^(41 x), each term exact in FP64; compensated sum
    // from the top limb down to limb 1, then the low limb
    double sm = t[NL - 1] * scale[NL - 1], err = 0.0;
#pragma unroll
    for (int x = NL - 2; x >= 1; --x) {
        const double h = t[x] * scale[x];
        const double sn = sm + h;
        const double bb = sn - sm;
        err += (sm - (sn - bb)) + (h - bb);
        sm = sn;
    }
    return sm + (err + t[0]);
}

}  // namespace cosma_amd
