// Split-precision ("Ozaki") DGEMM: FP64-accurate C <- alpha op(A) op(B) + beta C
// computed on the INT8 matrix cores.
//
// No counterpart in the reference (COSMA's local multiply is a plain host or
// vendor DGEMM); this is the split-precision path beside the FP64 MFMA DGEMM.
//
// Both methods start the same way (FP64, on the GPU):
//   1. exponent extraction: E_i = exponent of max_k |op(A)(i,k)| per row of
//      op(A), F_j per column of op(B) (frexp convention: max < 2^E);
//   2. power-of-two scaling of each row of A / column of B by 2^-E_i / 2^-F_j,
//      so every scaled entry lies in (-1, 1);
// then split the scaled entries into INT8 operands:
//   * moduli (default): the scaled entries are rounded to integers of P bits,
//     a' = rint(a * 2^(P - E_i)), and represented by their symmetric residues
//     modulo N pairwise-coprime moduli m_l <= 256 (each residue fits int8).
//     One INT8 GEMM per modulus gives A'B' mod m_l exactly (INT32
//     accumulation); the Chinese-remainder recombination of the N residues
//     recovers the exact integer A'B' (|A'B'| < M / 2, M = prod m_l), which is
//     scaled back by 2^(E_i + F_j - 2P).  The only approximation is the
//     rounding of the inputs to P bits (relative to their row / column max).
//   * slices: the scaled entries are written as sums of s signed 7-bit digits
//     (base 2^7, round to nearest, the leading digit within [-127, 127]);
//     the slice products with t + u <= s + 1 are accumulated exactly in INT32
//     per weight 2^-7(t+u) and recombined in FP64.
// FP64 is used on the GPU only for the exponent extraction and scaling, the
// recombination, and the alpha * AB + beta * C epilogue.
// The number of moduli (or slices) is chosen from the requested tolerance; the
// a-priori bound and the count used are returned to the caller.
#include <climits>
#include <cmath>

#include "i8gemm.hpp"

namespace cosma_amd {

// pairwise-coprime moduli, largest first (256 = 2^8; 255 = 3*5*17; 253 = 11*23;
// 247 = 13*19; 217 = 7*31; the rest are primes)
static const int kModuli[16] = {256, 255, 253, 251, 247, 241, 239, 233, 229, 227, 223, 217, 211, 199, 197, 193};
constexpr int MAX_MODULI = 16;
constexpr int LIMB = 42;   // CRT limb width in bits (3 limbs cover M < 2^126)

struct CrtConst {
    int n;
    int m[MAX_MODULI];
    double inv_m[MAX_MODULI];
    double w[MAX_MODULI][3];    // CRT weights W_l in 42-bit limbs (exact in FP64)
    int64_t mlimb[3];           // M in 42-bit limbs
    double inv_M;               // 1 / M
};

namespace {

// ------------------------------------------------------------ exponents
// maxbits[r] = bit pattern of max |op(X)(r, :)| (non-negative doubles order
// like their bit patterns, so an unsigned atomic max gives the numeric max;
// NaN patterns compare above +inf and are detected afterwards).
__global__ void clear_u64(unsigned long long *p, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0ull;
}

// rows of op(X) are strided in memory (X 'N' for A, X 'T' for B): thread per
// row, 64 columns per block-row of the grid.
__global__ void rowmax_strided(const double *__restrict__ x, int64_t ld, int64_t R, int64_t K,
                               unsigned long long *maxbits) {
    const int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (r >= R) return;
    const int64_t k0 = (int64_t)blockIdx.y * 64, k1 = min(K, k0 + 64);
    unsigned long long best = 0;
    for (int64_t k = k0; k < k1; ++k) {
        const unsigned long long b = __double_as_longlong(fabs(x[r + k * ld]));
        best = b > best ? b : best;
    }
    atomicMax(&maxbits[r], best);
}

// rows of op(X) are contiguous (X 'T' for A, 'N' for B): block per (row, chunk)
__global__ void rowmax_contig(const double *__restrict__ x, int64_t ld, int64_t R, int64_t K,
                              unsigned long long *maxbits) {
    const int64_t r = blockIdx.x;
    const int64_t k0 = (int64_t)blockIdx.y * 8192, k1 = min(K, k0 + 8192);
    unsigned long long best = 0;
    for (int64_t k = k0 + threadIdx.x; k < k1; k += blockDim.x) {
        const unsigned long long b = __double_as_longlong(fabs(x[k + r * ld]));
        best = b > best ? b : best;
    }
    __shared__ unsigned long long red[256];
    red[threadIdx.x] = best;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s && red[threadIdx.x + s] > red[threadIdx.x]) red[threadIdx.x] = red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicMax(&maxbits[r], red[0]);
}

// Row exponent from the max: max < 2^E.  For slices the leading digit must
// stay within [-127, 127]: rows whose scaled max would round to 128 take E + 1.
__device__ __forceinline__ int row_exponent(unsigned long long bits, int slices) {
    const double mx = __longlong_as_double((long long)bits);
    if (mx == 0.0) return 0;
    int e;
    const double f = frexp(mx, &e);   // mx = f * 2^e, f in [0.5, 1)
    if (slices && f * 128.0 >= 127.5) ++e;
    return e;
}

__global__ void exponents_kernel(const unsigned long long *maxbits, int64_t R, int slices, int *E,
                                 int *nonfinite) {
    for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < R; r += (int64_t)gridDim.x * blockDim.x) {
        const unsigned long long b = maxbits[r];
        if (b >= 0x7ff0000000000000ull) atomicOr(nonfinite, 1);   // inf or NaN
        E[r] = row_exponent(b, slices);
    }
}

// ----------------------------------------------------------------- split
struct SplitArgs {
    const double *x;
    int64_t ld;
    int64_t R, K;          // extents of op(X) (rows, shared dim)
    int64_t Kpad;          // row stride of the int8 output (bytes)
    int64_t slice_stride;  // bytes between slices / residues
    const int *E;
    int method;            // 0 moduli, 1 slices
    int count;             // number of moduli / slices
    int P;                 // bits of the integer inputs (moduli)
    int m[MAX_MODULI];
    double inv_m[MAX_MODULI];
    int8_t *out;
};

// 16 consecutive k of one row -> 16 bytes per modulus / slice
__device__ __forceinline__ void split16(const SplitArgs &a, const double (&v)[16], int e, int8_t *dst) {
    if (a.method == 0) {
        double iv[16];
#pragma unroll
        for (int j = 0; j < 16; ++j) iv[j] = rint(ldexp(v[j], a.P - e));
        for (int l = 0; l < a.count; ++l) {
            const double m = (double)a.m[l], im = a.inv_m[l];
            const int hi = (a.m[l] - 1) / 2, lo = -(a.m[l] / 2);
            union { i32x4 q; int8_t b[16]; } u;
#pragma unroll
            for (int j = 0; j < 16; ++j) {
                int r = (int)fma(-m, rint(iv[j] * im), iv[j]);   // exact: integers below 2^53
                if (r > hi) r -= a.m[l];
                if (r < lo) r += a.m[l];
                u.b[j] = (int8_t)r;
            }
            *reinterpret_cast<i32x4 *>(dst + (int64_t)l * a.slice_stride) = u.q;
        }
    } else {
        double r[16];
#pragma unroll
        for (int j = 0; j < 16; ++j) r[j] = ldexp(v[j], -e);
        for (int t = 0; t < a.count; ++t) {
            union { i32x4 q; int8_t b[16]; } u;
#pragma unroll
            for (int j = 0; j < 16; ++j) {
                const double y = r[j] * 128.0;     // exact
                const double d = rint(y);
                r[j] = y - d;                      // exact, |r| <= 1/2
                u.b[j] = (int8_t)(int)d;
            }
            *reinterpret_cast<i32x4 *>(dst + (int64_t)t * a.slice_stride) = u.q;
        }
    }
}

// op(X) rows contiguous in k (A 'T', B 'N'): thread per (row, 16-k chunk).
// Rows in [R, Rpad) and k in [K, Kpad) are written as zeros.
__global__ void split_contig(SplitArgs a, int64_t Rpad) {
    const int64_t chunks = a.Kpad / 16;
    const int64_t total = Rpad * chunks;
    for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < total;
         idx += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = idx / chunks, k0 = (idx - r * chunks) * 16;
        double v[16];
        const bool rin = r < a.R;
        if (rin && k0 + 16 <= a.K && ((a.ld & 1) == 0) && ((reinterpret_cast<uintptr_t>(a.x) & 15) == 0)) {
            const f64x2 *p = reinterpret_cast<const f64x2 *>(a.x + k0 + r * a.ld);
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const f64x2 w = p[j];
                v[2 * j] = w.x;
                v[2 * j + 1] = w.y;
            }
        } else {
#pragma unroll
            for (int j = 0; j < 16; ++j) v[j] = (rin && k0 + j < a.K) ? a.x[k0 + j + r * a.ld] : 0.0;
        }
        split16(a, v, rin ? a.E[r] : 0, a.out + r * a.Kpad + k0);
    }
}

// op(X) rows strided (A 'N', B 'T'): 64 x 64 tiles transposed through LDS.
// blockIdx.x -> 64 rows, blockIdx.y -> 64 k.
__global__ __launch_bounds__(256) void split_strided(SplitArgs a) {
    __shared__ double tile[64][65];
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    const int64_t r0 = (int64_t)blockIdx.x * 64, k0 = (int64_t)blockIdx.y * 64;
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int kk = ty + 4 * q;
        const int64_t r = r0 + tx, k = k0 + kk;
        tile[kk][tx] = (r < a.R && k < a.K) ? a.x[r + k * a.ld] : 0.0;
    }
    __syncthreads();
    const int row = threadIdx.x >> 2, kc = threadIdx.x & 3;
    double v[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) v[j] = tile[kc * 16 + j][row];
    const int64_t r = r0 + row;
    split16(a, v, r < a.R ? a.E[r] : 0, a.out + r * a.Kpad + k0 + kc * 16);
}

// ------------------------------------------------------ recombination
// Each thread recombines 4 consecutive rows of one column (one 4-byte load
// per residue plane; ldr is a multiple of 4).
__global__ void crt_epilogue(CrtConst cc, const int8_t *__restrict__ res, int64_t ldr, int64_t res_stride,
                             int64_t M, int64_t N, int P, const int *__restrict__ E, const int *__restrict__ F,
                             double alpha, double beta, double *__restrict__ C, int64_t ldc) {
    const int64_t quads = ldr / 4;
    const int64_t total = quads * N;
    for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < total;
         idx += (int64_t)gridDim.x * blockDim.x) {
        const int64_t j = idx / quads, i0 = (idx - j * quads) * 4;
        if (i0 >= M) continue;
        const int8_t *rp = res + i0 + j * ldr;
        // S = sum_l c_l W_l limb by limb: |c_l| <= 128 and limbs < 2^42, so
        // every partial sum of at most 16 terms stays below 2^53 (exact FMAs)
        double s[4][3];
#pragma unroll
        for (int e = 0; e < 4; ++e) s[e][0] = s[e][1] = s[e][2] = 0.0;
        for (int l = 0; l < cc.n; ++l) {
            const uint32_t word = *reinterpret_cast<const uint32_t *>(rp + (int64_t)l * res_stride);
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const double c = (double)(int8_t)(word >> (8 * e));
                s[e][0] = fma(c, cc.w[l][0], s[e][0]);
                s[e][1] = fma(c, cc.w[l][1], s[e][1]);
                s[e][2] = fma(c, cc.w[l][2], s[e][2]);
            }
        }
        const double fj = (double)F[j];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int64_t i = i0 + e;
            if (i >= M) break;
            // q = round(S / M); |A'B'| < M / 2 by the choice of P, and sd * inv_M
            // is within ~2^-40 of S / M, so the rounding cannot fall on the
            // wrong side of a tie
            const double sd = ldexp(s[e][2], 2 * LIMB) + ldexp(s[e][1], LIMB) + s[e][0];
            const int64_t q = (int64_t)rint(sd * cc.inv_M);
            int64_t t0 = (int64_t)s[e][0] - q * cc.mlimb[0], t1 = (int64_t)s[e][1] - q * cc.mlimb[1];
            int64_t t2 = (int64_t)s[e][2] - q * cc.mlimb[2];
            int64_t carry = t0 >> LIMB;              // arithmetic shift: floor division
            t0 -= carry << LIMB;
            t1 += carry;
            carry = t1 >> LIMB;
            t1 -= carry << LIMB;
            t2 += carry;
            // value = t2 * 2^84 + t1 * 2^42 + t0, limbs exact in double
            const double hi = ldexp((double)t2, 2 * LIMB), mid = ldexp((double)t1, LIMB);
            const double sm = hi + mid;
            const double bb = sm - hi;
            const double err = (hi - (sm - bb)) + (mid - bb);   // TwoSum error of hi + mid
            const double v = sm + (err + (double)t0);
            const double val = alpha * ldexp(v, E[i] + (int)fj - 2 * P);
            double *cp = C + i + j * ldc;
            *cp = (beta == 0.0) ? val : fma(beta, *cp, val);
        }
    }
}

__global__ void slices_epilogue(const double *__restrict__ acc, int64_t lda, int64_t M, int64_t N,
                                const int *__restrict__ E, const int *__restrict__ F, double alpha, double beta,
                                double *__restrict__ C, int64_t ldc) {
    const int64_t total = M * N;
    for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < total;
         idx += (int64_t)gridDim.x * blockDim.x) {
        const int64_t j = idx / M, i = idx - j * M;
        const double val = alpha * ldexp(acc[i + j * lda], E[i] + F[j]);
        double *cp = C + i + j * ldc;
        *cp = (beta == 0.0) ? val : fma(beta, *cp, val);
    }
}

inline int64_t round_up(int64_t x, int64_t a) { return (x + a - 1) / a * a; }

int grid_for(int64_t work) {
    int64_t b = (work + 255) / 256;
    return (int)(b > 65536 ? 65536 : (b < 1 ? 1 : b));
}

}  // namespace

// ---------------------------------------------------------------- host side
double log2_moduli(int n) {
    double s = 0;
    for (int l = 0; l < n; ++l) s += std::log2((double)kModuli[l]);
    return s;
}

// Largest P with k * 2^(2P) < M / 2 (with a relative margin of ~1e-6): the
// integer inputs satisfy |a'| <= 2^P, so |A'B'| <= k * 2^(2P) stays inside
// the symmetric residue range (-M/2, M/2).  The recombination estimates the
// CRT quotient from an FP64 value of S / M that is accurate to ~2^-40, far
// inside that margin.
int moduli_bits(int n, int64_t k) {
    const double lm = log2_moduli(n);
    return (int)std::floor((lm - 1.0 - std::log2((double)std::max<int64_t>(k, 1)) - 1e-6) / 2.0);
}

static void crt_constants(int n, CrtConst &cc) {
    typedef unsigned __int128 u128;
    cc.n = n;
    u128 M = 1;
    for (int l = 0; l < n; ++l) M *= (u128)kModuli[l];
    const u128 mask = ((u128)1 << LIMB) - 1;
    for (int l = 0; l < n; ++l) {
        const int m = kModuli[l];
        cc.m[l] = m;
        cc.inv_m[l] = 1.0 / m;
        const u128 Ml = M / (u128)m;
        const int r = (int)(Ml % (u128)m);
        int inv = 1;
        for (int y = 1; y < m; ++y)
            if ((r * y) % m == 1) { inv = y; break; }
        const u128 W = (Ml * (u128)inv) % M;
        cc.w[l][0] = (double)(int64_t)(W & mask);
        cc.w[l][1] = (double)(int64_t)((W >> LIMB) & mask);
        cc.w[l][2] = (double)(int64_t)(W >> (2 * LIMB));
    }
    cc.mlimb[0] = (int64_t)(M & mask);
    cc.mlimb[1] = (int64_t)((M >> LIMB) & mask);
    cc.mlimb[2] = (int64_t)(M >> (2 * LIMB));
    cc.inv_M = 1.0 / (double)M;
}

int ozaki_dgemm(cosma_amd_handle_t h, char ta, char tb, int64_t m, int64_t n, int64_t k, double alpha,
                const double *a, int64_t lda, const double *b, int64_t ldb, double beta, double *c,
                int64_t ldc, int method, double tol, int count, cosma_amd_ozaki_info *info, hipStream_t s) {
    if (info) *info = cosma_amd_ozaki_info{};
    if (m == 0 || n == 0) return COSMA_AMD_OK;
    if (k == 0 || alpha == 0.0) return launch_scale(m, n, beta, c, ldc, s);
    if (tol <= 0.0) tol = 1e-14;
    const double want_bits = std::log2(2.0 / tol);          // error model: 2 * 2^-P
    int P = 0;
    if (method == COSMA_AMD_OZAKI_MODULI) {
        if (count <= 0) {
            count = 2;
            while (count < MAX_MODULI && moduli_bits(count, k) < want_bits) ++count;
        }
        if (count < 2 || count > MAX_MODULI) {
            set_error("ozaki: %d moduli requested (2..%d supported)", count, MAX_MODULI);
            return COSMA_AMD_ERR_INVALID_ARG;
        }
        // residues stay exact while a' / m_l < 2^52 (one correction step in
        // the split) and a' fits int64: P <= 60
        P = std::min(moduli_bits(count, k), 60);
        if (P < 1) {
            set_error("ozaki: %d moduli leave no precision for k = %lld", count, (long long)k);
            return COSMA_AMD_ERR_UNSUPPORTED;
        }
    } else {
        if (count <= 0) count = (int)std::ceil((want_bits + 1.0) / 7.0);
        count = std::max(count, 1);
        if (count > I8_MAX_PAIRS) {
            set_error("ozaki: %d slices requested (1..%d supported)", count, I8_MAX_PAIRS);
            return COSMA_AMD_ERR_INVALID_ARG;
        }
        // INT32 exactness of one weight: pairs * k * 127^2 < 2^31
        if ((double)count * (double)k * 16129.0 >= 2147483648.0) {
            set_error("ozaki slices: k = %lld too large for exact INT32 accumulation", (long long)k);
            return COSMA_AMD_ERR_UNSUPPORTED;
        }
    }
    if (method == COSMA_AMD_OZAKI_MODULI && (double)k * 16384.0 >= 2147483648.0) {
        set_error("ozaki moduli: k = %lld too large for exact INT32 accumulation", (long long)k);
        return COSMA_AMD_ERR_UNSUPPORTED;
    }

    // ---- workspace
    const int64_t Mp = round_up(m, 256), Np = round_up(n, 256), Kp = round_up(k, 256);
    const int64_t ldr = round_up(m, 4);                         // residue output ld
    const size_t a_bytes = (size_t)count * Mp * Kp, b_bytes = (size_t)count * Np * Kp;
    const size_t out_bytes = method == COSMA_AMD_OZAKI_MODULI ? (size_t)count * ldr * n
                                                               : (size_t)ldr * n * 8;
    const size_t aux = (size_t)(Mp + Np + 1) * 8 + (size_t)(Mp + Np) * 4 + 256;
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    const size_t need = al(a_bytes) + al(b_bytes) + al(out_bytes) + al(aux);
    if (h->ozaki_ws.bytes < need) {
        if (h->ozaki_ws.ptr) CA_HIP(hipFree(h->ozaki_ws.ptr));
        h->ozaki_ws.ptr = nullptr;
        h->ozaki_ws.bytes = 0;
        hipError_t e = hipMalloc(&h->ozaki_ws.ptr, need);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            set_error("ozaki: workspace allocation of %zu bytes failed", need);
            return COSMA_AMD_ERR_ALLOC;
        }
        h->ozaki_ws.bytes = need;
    }
    char *ws = static_cast<char *>(h->ozaki_ws.ptr);
    int8_t *As = reinterpret_cast<int8_t *>(ws);
    int8_t *Bs = reinterpret_cast<int8_t *>(ws + al(a_bytes));
    void *out = ws + al(a_bytes) + al(b_bytes);
    char *auxp = ws + al(a_bytes) + al(b_bytes) + al(out_bytes);
    // aux: max |A| per row [Mp], max |B| per column [Np], non-finite flag [1]
    // (all cleared together), then the exponents E [Mp] and F [Np]
    unsigned long long *maxA = reinterpret_cast<unsigned long long *>(auxp);
    unsigned long long *maxB = maxA + Mp;
    int *flag = reinterpret_cast<int *>(maxB + Np);
    int *E = reinterpret_cast<int *>(maxB + Np + 1);
    int *F = E + Mp;

    // ---- 1. exponents
    hipLaunchKernelGGL(clear_u64, dim3(grid_for(Mp + Np + 1)), dim3(256), 0, s, maxA, Mp + Np + 1);
    const bool a_rows_contig = ta == 'T', b_rows_contig = tb == 'N';   // op rows contiguous in k
    if (a_rows_contig)
        hipLaunchKernelGGL(rowmax_contig, dim3((unsigned)m, (unsigned)((k + 8191) / 8192)), dim3(256), 0, s, a, lda, m, k, maxA);
    else
        hipLaunchKernelGGL(rowmax_strided, dim3((unsigned)((m + 255) / 256), (unsigned)((k + 63) / 64)), dim3(256), 0, s, a, lda, m, k, maxA);
    if (b_rows_contig)
        hipLaunchKernelGGL(rowmax_contig, dim3((unsigned)n, (unsigned)((k + 8191) / 8192)), dim3(256), 0, s, b, ldb, n, k, maxB);
    else
        hipLaunchKernelGGL(rowmax_strided, dim3((unsigned)((n + 255) / 256), (unsigned)((k + 63) / 64)), dim3(256), 0, s, b, ldb, n, k, maxB);
    const int slices = method == COSMA_AMD_OZAKI_SLICES;
    hipLaunchKernelGGL(exponents_kernel, dim3(grid_for(m)), dim3(256), 0, s, maxA, m, slices, E, flag);
    hipLaunchKernelGGL(exponents_kernel, dim3(grid_for(n)), dim3(256), 0, s, maxB, n, slices, F, flag);
    CA_HIP(hipGetLastError());
    int nonfinite = 0;
    CA_HIP(hipMemcpyAsync(&nonfinite, flag, sizeof(int), hipMemcpyDeviceToHost, s));
    CA_HIP(hipStreamSynchronize(s));
    if (nonfinite) {
        // Inf / NaN cannot be split; the FP64 MFMA DGEMM propagates them as IEEE does.
        if (info) info->method = -1;
        return launch_dgemm(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, s);
    }

    // ---- 2. split
    auto split = [&](const double *x, int64_t ld, int64_t R, int64_t Rpad, bool contig, const int *ex,
                     int8_t *dst, int64_t stride) {
        SplitArgs sa{};
        sa.x = x; sa.ld = ld; sa.R = R; sa.K = k; sa.Kpad = Kp; sa.slice_stride = stride;
        sa.E = ex; sa.method = slices; sa.count = count; sa.P = P; sa.out = dst;
        for (int l = 0; l < MAX_MODULI; ++l) { sa.m[l] = kModuli[l]; sa.inv_m[l] = 1.0 / kModuli[l]; }
        if (contig)
            hipLaunchKernelGGL(split_contig, dim3(grid_for(Rpad * (Kp / 16))), dim3(256), 0, s, sa, Rpad);
        else
            hipLaunchKernelGGL(split_strided, dim3((unsigned)(Rpad / 64), (unsigned)(Kp / 64)), dim3(256), 0, s, sa);
    };
    split(a, lda, m, Mp, a_rows_contig, E, As, Mp * Kp);
    split(b, ldb, n, Np, b_rows_contig, F, Bs, Np * Kp);
    CA_HIP(hipGetLastError());

    // ---- 3. INT8 products
    I8Problem pr{};
    pr.M = m; pr.N = n; pr.K = Kp;
    pr.a = As; pr.lda = Kp; pr.a_stride = Mp * Kp;
    pr.b = Bs; pr.ldb = Kp; pr.b_stride = Np * Kp;
    pr.out = out; pr.ldo = ldr;
    I8Jobs jobs{};
    int rc;
    int64_t products = 0;
    if (!slices) {
        pr.out_stride = ldr * n;
        for (int l = 0; l < count; ++l) {
            jobs.npairs[l] = 1;
            jobs.pa[l][0] = (unsigned char)l;
            jobs.pb[l][0] = (unsigned char)l;
            jobs.modulus[l] = kModuli[l];
        }
        if ((rc = launch_i8gemm(pr, jobs, count, EPI_RES8, s))) return rc;
        products = count;
        CrtConst cc;
        crt_constants(count, cc);
        hipLaunchKernelGGL(crt_epilogue, dim3(grid_for(ldr / 4 * n)), dim3(256), 0, s, cc,
                           static_cast<const int8_t *>(out), ldr, ldr * n, m, n, P, E, F, alpha, beta, c, ldc);
    } else {
        pr.out_stride = 0;
        // weight d = t + u (slice indices from 1): pairs with t + u = d, d <= count + 1
        for (int d = 2; d <= count + 1; ++d) {
            I8Jobs jd{};
            int np = 0;
            for (int t = 1; t <= count; ++t) {
                const int u = d - t;
                if (u < 1 || u > count) continue;
                jd.pa[0][np] = (unsigned char)(t - 1);
                jd.pb[0][np] = (unsigned char)(u - 1);
                ++np;
            }
            jd.npairs[0] = np;
            jd.scale[0] = std::ldexp(1.0, -7 * d);
            pr.accumulate = d > 2;
            if ((rc = launch_i8gemm(pr, jd, 1, EPI_F64ACC, s))) return rc;
            products += np;
        }
        hipLaunchKernelGGL(slices_epilogue, dim3(grid_for(m * n)), dim3(256), 0, s,
                           static_cast<const double *>(out), ldr, m, n, E, F, alpha, beta, c, ldc);
    }
    CA_HIP(hipGetLastError());
    if (info) {
        info->method = slices ? COSMA_AMD_OZAKI_SLICES : COSMA_AMD_OZAKI_MODULI;
        info->count = count;
        info->bits = slices ? 7 * count : P;
        info->int8_products = products;
        // a-priori estimate of the relative Frobenius error for inputs whose
        // entries are comparable to their row / column maxima
        info->error_bound = slices ? 2.0 * count * std::ldexp(1.0, -7 * count) : 2.0 * std::ldexp(1.0, -P);
    }
    return COSMA_AMD_OK;
}

}  // namespace cosma_amd
