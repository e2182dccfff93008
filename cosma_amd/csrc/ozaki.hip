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
//     P is the largest value for which either the entrywise bound
//     |A'B'| <= k 2^(2P) or the row / column 2-norm bound (norm_bits) keeps
//     |A'B'| below M / 2.
//   * slices: the scaled entries are written as sums of s signed 7-bit digits
//     (base 2^7, round to nearest, the leading digit within [-127, 127]);
//     the slice products with t + u <= s + 1 are accumulated exactly in INT32
//     per weight 2^-7(t+u) and recombined in FP64.
// FP64 is used on the GPU only for the exponent extraction and scaling, the
// recombination, and the alpha * AB + beta * C epilogue.
// The number of moduli (or slices) is chosen from the requested tolerance; the
// a-priori bound and the count used are returned to the caller.
#include <algorithm>
#include <cfloat>
#include <climits>
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "ozaki.hpp"

namespace cosma_amd {

namespace {

// ------------------------------------------------------------ exponents
// maxbits[r] = bit pattern of max |op(X)(r, :)| (non-negative doubles order
// like their bit patterns, so an unsigned atomic max gives the numeric max;
// NaN patterns compare above +inf and are detected afterwards).
__global__ void clear_u64(unsigned long long *p, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0ull;
}

// Statistics of op(X) seen as R vectors of length K (the rows of op(A), the
// columns of op(B)): the max |x| of every vector (as bit patterns, for the
// exponents), the sum of squares of every vector (for the bound on |A'B'|)
// and, for every k, the sum of squares across the vectors (for the error
// estimate).  No LDS: a wave reads 64 consecutive elements along the
// contiguous dimension per load, keeps the reduction along the other
// dimension in registers, and reduces across its lanes with shuffles.
// CONTIG (element (r, k) at x[k + r * ld]): lane = k, wave = 64 vectors, one
// wave max and sum per vector and STATS_KC * 64 k; otherwise (x[r + k * ld]):
// lane = r within STATS_RCH chunks of 64 rows, wave = 64 k, one wave sum per
// k.  Workgroups cover STATS_KC * 64 (CONTIG) or STATS_RCH * 64 lanes x
// STATS_SPAN of the other dimension (4 waves x STATS_SPAN / 4).
// The contiguous form reads with non-temporal loads (8192^2: 140 -> 128 us).
constexpr int STATS_SPAN = 256;
constexpr int STATS_KC = 8;   // 64-wide k chunks per workgroup (contiguous vectors)
constexpr int STATS_RCH = 4;  // 64-row chunks per wave (strided vectors)

template <bool CONTIG>
__global__ __launch_bounds__(256) void stats_kernel(const double *__restrict__ x, int64_t ld, int64_t R, int64_t K,
                                                    unsigned long long *maxbits, double *ksq, double *rsq,
                                                    double *rabs) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    constexpr int PER_WAVE = STATS_SPAN / 4, U = 8;
    if (CONTIG) {
        // A workgroup covers STATS_KC chunks of 64 consecutive k; each lane
        // accumulates its k of every chunk for a group of U vectors before
        // the group's cross-lane reductions, so there is one reduction per
        // STATS_KC * 64 elements of a vector (and 8 x 8 loads in flight).
        const int64_t kbase = (int64_t)blockIdx.x * 64 * STATS_KC + lane;
        const int64_t rb = (int64_t)blockIdx.y * STATS_SPAN + wave * PER_WAVE;
        double ss[STATS_KC];
#pragma unroll
        for (int c = 0; c < STATS_KC; ++c) ss[c] = 0.0;
        for (int q0 = 0; q0 < PER_WAVE; q0 += U) {
            unsigned long long b[U];
            double q[U], l1[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                b[u] = 0;
                q[u] = 0.0;
                l1[u] = 0.0;
            }
#pragma unroll
            for (int c = 0; c < STATS_KC; ++c) {
                const int64_t k = kbase + 64 * c;
                double v[U];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const int64_t r = rb + q0 + u;
                    v[u] = (k < K && r < R) ? __builtin_nontemporal_load(x + k + r * ld) : 0.0;
                }
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    ss[c] = fma(v[u], v[u], ss[c]);
                    q[u] = fma(v[u], v[u], q[u]);
                    l1[u] += fabs(v[u]);
                    const unsigned long long bb = __double_as_longlong(fabs(v[u]));
                    b[u] = bb > b[u] ? bb : b[u];
                }
            }
            // U independent lane reductions (max and sum of squares), stepped
            // together so their shuffle latencies overlap
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const unsigned long long o = __shfl_xor(b[u], off);
                    b[u] = o > b[u] ? o : b[u];
                    q[u] += __shfl_xor(q[u], off);
                    l1[u] += __shfl_xor(l1[u], off);
                }
            }
            if (lane < U && rb + q0 + lane < R) {
                unsigned long long mine = b[0];
                double qs = q[0], as = l1[0];
#pragma unroll
                for (int u = 1; u < U; ++u) {
                    mine = lane == u ? b[u] : mine;
                    qs = lane == u ? q[u] : qs;
                    as = lane == u ? l1[u] : as;
                }
                atomicMax(&maxbits[rb + q0 + lane], mine);
                atomicAdd(&rsq[rb + q0 + lane], qs);
                atomicAdd(&rabs[rb + q0 + lane], as);
            }
        }
#pragma unroll
        for (int c = 0; c < STATS_KC; ++c)
            if (kbase + 64 * c < K) atomicAdd(&ksq[kbase + 64 * c], ss[c]);
    } else {
        // A wave covers STATS_RCH chunks of 64 rows (lane = row within a
        // chunk) for PER_WAVE k: the per-k sums of squares accumulate over
        // the chunks in the lane before one cross-lane reduction, and the
        // STATS_RCH * U loads of a group are all in flight together.
        const int64_t rbase = (int64_t)blockIdx.x * 64 * STATS_RCH + lane;
        const int64_t kb = (int64_t)blockIdx.y * STATS_SPAN + wave * PER_WAVE;
        unsigned long long best[STATS_RCH];
        double rs[STATS_RCH], ra[STATS_RCH];
#pragma unroll
        for (int c = 0; c < STATS_RCH; ++c) {
            best[c] = 0;
            rs[c] = 0.0;
            ra[c] = 0.0;
        }
        for (int q0 = 0; q0 < PER_WAVE; q0 += U) {
            double v[STATS_RCH][U];
#pragma unroll
            for (int c = 0; c < STATS_RCH; ++c) {
                const int64_t r = rbase + 64 * c;
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const int64_t k = kb + q0 + u;
                    v[c][u] = (r < R && k < K) ? x[r + k * ld] : 0.0;
                }
            }
            double sq[U];
#pragma unroll
            for (int u = 0; u < U; ++u) sq[u] = 0.0;
#pragma unroll
            for (int c = 0; c < STATS_RCH; ++c)
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const unsigned long long b = __double_as_longlong(fabs(v[c][u]));
                    best[c] = b > best[c] ? b : best[c];
                    const double q = v[c][u] * v[c][u];
                    sq[u] += q;
                    rs[c] += q;
                    ra[c] += fabs(v[c][u]);
                }
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
                for (int u = 0; u < U; ++u) sq[u] += __shfl_xor(sq[u], off);
            }
            if (lane < U && kb + q0 + lane < K) {
                double mine = sq[0];
#pragma unroll
                for (int u = 1; u < U; ++u) mine = lane == u ? sq[u] : mine;
                atomicAdd(&ksq[kb + q0 + lane], mine);
            }
        }
#pragma unroll
        for (int c = 0; c < STATS_RCH; ++c) {
            const int64_t r = rbase + 64 * c;
            if (r < R) {
                atomicMax(&maxbits[r], best[c]);
                atomicAdd(&rsq[r], rs[c]);
                atomicAdd(&rabs[r], ra[c]);
            }
        }
    }
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
    int64_t Kpad;          // padded k (elements)
    int64_t row_bytes;     // row stride of the output (bytes)
    int64_t slice_stride;  // bytes between slices / residues
    const int *E;
    int method;            // 0 moduli, 1 slices, 2 moduli with BF16 residues
    int esize;             // bytes per output element (2 for BF16)
    int count;             // number of moduli / slices
    int P;                 // bits of the integer inputs (moduli)
    int m[MAX_MODULI];
    double inv_m[MAX_MODULI];
    int8_t *out;
};

// 16 consecutive k of one row -> 16 bytes per modulus / slice
__device__ __forceinline__ void split16(const SplitArgs &a, const double (&v)[16], int e, int8_t *dst) {
    if (a.method == COSMA_AMD_OZAKI_MODULI) {
        moduli_residues16<false>(v, e, a.P, a.count, a.m, a.inv_m, dst, a.slice_stride);
    } else if (a.method == COSMA_AMD_OZAKI_MODULI_BF16) {
        moduli_residues16<true>(v, e, a.P, a.count, a.m, a.inv_m, dst, a.slice_stride);
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

// Order of k inside the residue / slice rows: within each block of 256 k,
// byte (element) 16 l + j of the row holds k = 32 (j >> 1) + 2 l + (j & 1)
// for l, j in [0, 16).  Both operands use the same order, so the INT8
// products over k are unchanged (exact integer sums), and the split kernels
// read whole 256-byte runs of the FP64 rows per 16 lanes (lane l loads the
// pairs (k, k + 1) at 32 q + 2 l) while each lane still writes its 16 values
// as one 16-byte run.
__device__ __forceinline__ int block_k(int l, int j) { return 32 * (j >> 1) + 2 * l + (j & 1); }

// op(X) rows contiguous in k (A 'T', B 'N'): 16 lanes per (row, 256-k block).
// Rows in [R, Rpad) and k in [K, Kpad) are written as zeros.
__global__ __launch_bounds__(256) void split_contig(SplitArgs a, int64_t Rpad) {
    const int64_t blocks = a.Kpad / 256;
    const int64_t total = Rpad * blocks * 16;
    const bool vec = ((a.ld & 1) == 0) && ((reinterpret_cast<uintptr_t>(a.x) & 15) == 0);
    for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < total;
         idx += (int64_t)gridDim.x * blockDim.x) {
        const int l = (int)(idx & 15);
        const int64_t rb = idx >> 4;
        const int64_t r = rb / blocks, k0 = (rb - r * blocks) * 256;
        const bool rin = r < a.R;
        const double *x = a.x + r * a.ld + k0;
        double v[16];
        if (rin && vec && k0 + 256 <= a.K) {
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const f64x2 w = *reinterpret_cast<const f64x2 *>(x + 32 * q + 2 * l);
                v[2 * q] = w.x;
                v[2 * q + 1] = w.y;
            }
        } else {
#pragma unroll
            for (int j = 0; j < 16; ++j) {
                const int kk = block_k(l, j);
                v[j] = (rin && k0 + kk < a.K) ? x[kk] : 0.0;
            }
        }
        split16(a, v, rin ? a.E[r] : 0, a.out + r * a.row_bytes + (k0 + 16 * l) * a.esize);
    }
}

// op(X) rows strided (A 'N', B 'T'): 16 rows x 256 k per workgroup,
// transposed through LDS (non-temporal loads).  Each k of the tile is one
// 128-byte run of the 16 rows; each output row leaves as one 256-byte run
// (16 lanes) per residue plane.
// blockIdx.x -> 16 rows, blockIdx.y -> 256 k.
__global__ __launch_bounds__(256) void split_strided(SplitArgs a) {
    __shared__ double tile[256][17];
    const int t = threadIdx.x;
    const int64_t r0 = (int64_t)blockIdx.x * 16, k0 = (int64_t)blockIdx.y * 256;
    {
        const int row = t & 15, kk = t >> 4;
        const int64_t r = r0 + row;
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int k = kk + 16 * q;
            tile[k][row] = (r < a.R && k0 + k < a.K) ? __builtin_nontemporal_load(a.x + r + (k0 + k) * a.ld) : 0.0;
        }
    }
    __syncthreads();
    const int l = t & 15, row = t >> 4;
    double v[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) v[j] = tile[block_k(l, j)][row];
    const int64_t r = r0 + row;
    split16(a, v, r < a.R ? a.E[r] : 0, a.out + r * a.row_bytes + (k0 + 16 * l) * a.esize);
}

// ------------------------------------------------------ recombination
// Each thread recombines 4 consecutive rows of one column (one 4-byte load
// per residue plane, the plane loop unrolled by 4 so that several loads are
// in flight; 8 rows per thread or no unrolling measured 1.3x-1.6x slower);
// blockIdx.y strides over the columns.  The limb sums and crt_value are exact
// FP64 integer arithmetic (see crt_value).
constexpr int CRT_ROWS = 4;
constexpr int64_t PLANE_PAD = 3 * 1024;   // bytes; a multiple of 16, not of a large power of two
// The epilogues also return the Frobenius norm of the product X = op(A)op(B)
// as computed (before alpha and beta), for the check of the error against
// the tolerance: sum of (X_ij 2^-shift)^2 in NORM_SLOTS partial sums (one
// atomic add per workgroup).  Epilogue modes:
constexpr int EPI_WRITE = 0;        // C <- alpha X + beta C
constexpr int EPI_WRITE_NORM = 1;   // the same, and the norm of X
constexpr int EPI_NORM = 2;         // the norm of X only (C untouched)
constexpr int NORM_SLOTS = 64;

// Adds v (one per thread) over the workgroup into nrm[slot] (all threads of
// the workgroup must call it).
__device__ __forceinline__ void block_norm_add(double v, double *nrm) {
    __shared__ double part[8];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0) part[wave] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += part[w];
        atomicAdd(&nrm[(blockIdx.x + blockIdx.y * gridDim.x) % NORM_SLOTS], t);
    }
}

// NL limbs of 41 bits: 3 while M < 2^122 (up to 15 moduli), else 4.
template <int NL, int MODE>
__global__ __launch_bounds__(256) void crt_epilogue(CrtConst cc, const int8_t *__restrict__ res, int64_t ldr,
                                                    int64_t res_stride, int64_t M, int64_t N, int P,
                                                    const int *__restrict__ E, const int *__restrict__ F,
                                                    double alpha, double beta, double *__restrict__ C,
                                                    int64_t ldc, int shift, double *__restrict__ nrm) {
    const int64_t i0 = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) * CRT_ROWS;
    double ss = 0.0;
    for (int64_t j = blockIdx.y; i0 < M && j < N; j += gridDim.y) {
        const int8_t *rp = res + i0 + j * ldr;
        // S = sum_l c_l W_l limb by limb
        double s[CRT_ROWS][NL];
#pragma unroll
        for (int e = 0; e < CRT_ROWS; ++e)
#pragma unroll
            for (int q = 0; q < NL; ++q) s[e][q] = 0.0;
#pragma unroll 4
        for (int l = 0; l < cc.n; ++l) {
            const uint32_t word = *reinterpret_cast<const uint32_t *>(rp + (int64_t)l * res_stride);
#pragma unroll
            for (int e = 0; e < CRT_ROWS; ++e) {
                const double c = (double)(int8_t)(word >> (8 * e));
#pragma unroll
                for (int q = 0; q < NL; ++q) s[e][q] = fma(c, cc.w[l][q], s[e][q]);
            }
        }
        const int fj = F[j];
#pragma unroll
        for (int e = 0; e < CRT_ROWS; ++e) {
            const int64_t i = i0 + e;
            if (i >= M) break;
            const double v = crt_value<NL>(s[e], cc);
            if (MODE != EPI_WRITE) {
                const double x = ldexp(v, E[i] + fj - 2 * P - shift);
                ss = fma(x, x, ss);
            }
            if (MODE != EPI_NORM) {
                const double val = alpha * ldexp(v, E[i] + fj - 2 * P);
                double *cp = C + i + j * ldc;
                *cp = (beta == 0.0) ? val : fma(beta, *cp, val);
            }
        }
    }
    if (MODE != EPI_WRITE) block_norm_add(ss, nrm);
}

template <int MODE>
__global__ __launch_bounds__(256) void slices_epilogue(const double *__restrict__ acc, int64_t lda, int64_t M,
                                                       int64_t N, const int *__restrict__ E,
                                                       const int *__restrict__ F, double alpha, double beta,
                                                       double *__restrict__ C, int64_t ldc, int shift,
                                                       double *__restrict__ nrm) {
    const int64_t total = M * N;
    double ss = 0.0;
    for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < total;
         idx += (int64_t)gridDim.x * blockDim.x) {
        const int64_t j = idx / M, i = idx - j * M;
        const double v = acc[i + j * lda];
        if (MODE != EPI_WRITE) {
            const double x = ldexp(v, E[i] + F[j] - shift);
            ss = fma(x, x, ss);
        }
        if (MODE != EPI_NORM) {
            const double val = alpha * ldexp(v, E[i] + F[j]);
            double *cp = C + i + j * ldc;
            *cp = (beta == 0.0) ? val : fma(beta, *cp, val);
        }
    }
    if (MODE != EPI_WRITE) block_norm_add(ss, nrm);
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

double moduli_half_limit(int n) { return std::exp2(log2_moduli(n) - 1.0 - 1e-6); }

namespace {
// Little-endian 256-bit unsigned integer, enough for M = prod of up to 20
// moduli (< 2^160) and the CRT weights W_l < M.
struct U256 {
    uint64_t w[4] = {0, 0, 0, 0};
};

void mul_small(U256 &x, uint32_t v) {
    unsigned __int128 carry = 0;
    for (int i = 0; i < 4; ++i) {
        const unsigned __int128 t = (unsigned __int128)x.w[i] * v + carry;
        x.w[i] = (uint64_t)t;
        carry = t >> 64;
    }
}

uint32_t div_small(U256 &x, uint32_t v) {   // x /= v, returns x % v
    unsigned __int128 rem = 0;
    for (int i = 3; i >= 0; --i) {
        const unsigned __int128 cur = (rem << 64) | x.w[i];
        x.w[i] = (uint64_t)(cur / v);
        rem = cur % v;
    }
    return (uint32_t)rem;
}

uint64_t bit_field(const U256 &x, int lo, int n) {   // bits [lo, lo + n), n <= 63
    uint64_t out = 0;
    for (int b = 0; b < n; ++b) {
        const int bit = lo + b;
        if (bit < 256 && ((x.w[bit / 64] >> (bit % 64)) & 1)) out |= (uint64_t)1 << b;
    }
    return out;
}

double to_double(const U256 &x) {
    double d = 0.0;
    for (int i = 3; i >= 0; --i) d = d * 18446744073709551616.0 + (double)x.w[i];
    return d;
}
}  // namespace

void crt_constants(int n, CrtConst &cc) {
    cc.n = n;
    U256 M;
    M.w[0] = 1;
    for (int l = 0; l < n; ++l) mul_small(M, (uint32_t)kModuli[l]);
    for (int l = 0; l < n; ++l) {
        const int m = kModuli[l];
        cc.m[l] = m;
        cc.inv_m[l] = 1.0 / m;
        U256 Ml = M;
        div_small(Ml, (uint32_t)m);                  // M / m_l, exact
        U256 tmp = Ml;
        const int r = (int)div_small(tmp, (uint32_t)m);   // (M / m_l) mod m_l
        int inv = 1;
        for (int y = 1; y < m; ++y)
            if ((r * y) % m == 1) { inv = y; break; }
        U256 W = Ml;
        mul_small(W, (uint32_t)inv);                 // W_l = (M / m_l) * inv < M
        for (int q = 0; q < NLIMB; ++q) cc.w[l][q] = (double)bit_field(W, q * LIMB, LIMB);
    }
    for (int q = 0; q < NLIMB; ++q) cc.mlimb[q] = (int64_t)bit_field(M, q * LIMB, LIMB);
    cc.inv_M = 1.0 / to_double(M);
}

namespace {
// Error model of the split-precision product X' (what the epilogue computes
// before alpha and beta) against X = op(A) op(B).  Rounding the entries of
// row i of op(A) to the grid of spacing 2^(E_i - P) moves each by at most
// delta_i = 2^(E_i - P - 1) (columns of op(B): eps_j = 2^(F_j - P - 1)), and
//   X' - X = dA B + A dB + dA dB,
// so, entry by entry,
//   |X'_ij - X_ij| <= delta_i ||op(B)(:, j)||_1 + eps_j ||op(A)(i, :)||_1 + k delta_i eps_j,
// whose Frobenius norm is at most
//   ||delta|| ||b1|| + ||a1|| ||eps|| + k ||delta|| ||eps||     (a rigorous bound),
// a1_i, b1_j the 1-norms of the rows of op(A) / columns of op(B).  With
// independent rounding errors of variance delta^2 / 3 the expected square is
//   (1/3) (||delta||^2 ||B||_F^2 + ||A||_F^2 ||eps||^2)           (the estimate).
// Slices: s digits of 7 bits leave delta_i = 2^(E_i - 7s - 1); the products of
// digit pairs with t + u >= s + 2 that are not computed add at most
// k D_s 2^(E_i + F_j), D_s = sum_{d = s+2}^{2s} (2s + 1 - d) 64^2 2^-7d (the
// digits after the leading one lie in [-64, 64]).
// Everything is held relative to 2^(ea + fb), ea / fb the largest exponents.
struct ErrModel {
    bool zero = false;      // a zero operand (every row or column exponent is the zero sentinel)
    int ea = 0, fb = 0;
    int64_t k = 0;
    double sA = 0, sB = 0;  // ||2^(E - ea)||, ||2^(F - fb)||
    double a1 = 0, b1 = 0;  // ||a1 2^-ea||, ||b1 2^-fb||
    double fA = 0, fB = 0;  // ||op(A)||_F 2^-ea, ||op(B)||_F 2^-fb
    double d0 = 0;          // sqrt(sum_l ||op(A)(:, l)||^2 ||op(B)(l, :)||^2) 2^-(ea + fb): the
                            // size of X the a-priori estimate assumes (independent signs)
    double gamma = 0;       // a-priori relative estimate = gamma * 2^-P (doubled where used)
};

// Powers of two 2^e for e in [-1100, 1099] (0 or inf outside the double range).
const double *pow2_table() {
    static const std::vector<double> t = [] {
        std::vector<double> v(2200);
        for (int e = -1100; e < 1100; ++e) v[e + 1100] = std::ldexp(1.0, e);
        return v;
    }();
    return t.data() + 1100;
}

// std::ldexp(x, e) without its per-call cost (tens of thousands of calls per
// product made the host statistics take ~0.4 ms at 8192^3): while 2^e is a
// normal double, x * 2^e is the correctly rounded value of x 2^e (a single
// rounding, also for subnormal or overflowing results), which is what ldexp
// returns; ldexp itself otherwise.  (P2 = pow2_table(), fetched once outside
// the loops.)
inline double scale2(const double *P2, double x, int e) {
    return (e >= -1022 && e <= 1023) ? x * P2[e] : std::ldexp(x, e);
}

// 2^e for e <= 0: the table holds ldexp(1, e) down to -1100, and every
// smaller power rounds to zero.
inline double pow2_nonpos(const double *P2, int e) { return e >= -1100 ? P2[e] : 0.0; }

// A read-only view of n host values.
template <typename T>
struct View {
    const T *p;
    int64_t n;
    const T &operator[](int64_t i) const { return p[i]; }
    int64_t size() const { return n; }
    const T *begin() const { return p; }
    const T *end() const { return p + n; }
};

ErrModel error_model(View<int> E, View<int> F, View<double> ksqA, View<double> ksqB, View<double> l1A,
                     View<double> l1B, int64_t k) {
    ErrModel em;
    em.k = k;
    int ea = INT_MIN, fb = INT_MIN;
    for (int e : E) if (e != ZERO_EXPONENT) ea = std::max(ea, e);
    for (int f : F) if (f != ZERO_EXPONENT) fb = std::max(fb, f);
    if (ea == INT_MIN || fb == INT_MIN) {
        em.zero = true;
        return em;
    }
    em.ea = ea;
    em.fb = fb;
    const double *P2 = pow2_table();
    double SA = 0.0, SB = 0.0, NA = 0.0, NB = 0.0, D = 0.0, LA = 0.0, LB = 0.0;
    // the same sums as with scale2 per term: the loop-invariant powers are
    // hoisted (a product by an exact normal power of two is scale2's result)
    auto vec_sums = [P2](View<int> ex, View<double> l1, int top, double &S_out, double &L_out) {
        const bool normal = -top >= -1022 && -top <= 1023;
        const double s = normal ? P2[-top] : 0.0;
        const int *e = ex.p;
        const double *x = l1.p;
        double S = 0.0, L = 0.0;
        for (int64_t i = 0; i < ex.size(); ++i) {
            if (e[i] == ZERO_EXPONENT) continue;
            S += pow2_nonpos(P2, 2 * (e[i] - top));
            const double a = normal ? x[i] * s : std::ldexp(x[i], -top);
            L += a * a;
        }
        S_out = S;
        L_out = L;
    };
    vec_sums(E, l1A, ea, SA, LA);
    vec_sums(F, l1B, fb, SB, LB);
    bool finite = true;
    {
        const bool na = -2 * ea >= -1022 && -2 * ea <= 1023, nb = -2 * fb >= -1022 && -2 * fb <= 1023;
        const double sa = na ? P2[-2 * ea] : 0.0, sb = nb ? P2[-2 * fb] : 0.0;
        const double *qa = ksqA.p, *qb = ksqB.p;
        double na_ = 0.0, nb_ = 0.0, d_ = 0.0;
        int nonfin = 0;
        for (int64_t l = 0; l < ksqA.size(); ++l) {
            nonfin |= !std::isfinite(qa[l]) | !std::isfinite(qb[l]);
            const double a = na ? qa[l] * sa : std::ldexp(qa[l], -2 * ea);
            const double b = nb ? qb[l] * sb : std::ldexp(qb[l], -2 * fb);
            na_ += a;
            nb_ += b;
            d_ += a * b;
        }
        NA = na_;
        NB = nb_;
        D = d_;
        finite = !nonfin;
    }
    // the 1-norms are FP64 sums of |x|: a relative margin covers their rounding
    const double margin = 1.0 + 1e-12;
    em.sA = std::sqrt(SA);
    em.sB = std::sqrt(SB);
    em.a1 = std::sqrt(LA) * margin;
    em.b1 = std::sqrt(LB) * margin;
    em.fA = std::sqrt(NA);
    em.fB = std::sqrt(NB);
    em.d0 = std::sqrt(D);
    em.gamma = (finite && D > 0.0) ? std::sqrt((SA * NB + NA * SB) / (12.0 * D)) : std::ldexp(1.0, 30);
    if (!std::isfinite(em.gamma)) em.gamma = std::ldexp(1.0, 30);
    return em;
}

// Rigorous bound on ||X' - X||_F relative to 2^(ea + fb) for P bits (moduli)
// or s slices.
double rigorous_num(const ErrModel &em, bool slices, int P_or_s) {
    if (em.zero) return 0.0;
    const double kk = (double)em.k;
    if (!slices) {
        const double d = std::ldexp(1.0, -P_or_s - 1);
        return d * (em.sA * em.b1 + em.a1 * em.sB) + kk * d * d * em.sA * em.sB;
    }
    const int sl = P_or_s;
    const double d = std::ldexp(1.0, -7 * sl - 1);
    double Ds = 0.0;
    for (int dd = sl + 2; dd <= 2 * sl; ++dd) Ds += (double)(2 * sl + 1 - dd) * 4096.0 * std::ldexp(1.0, -7 * dd);
    return d * (em.sA * em.b1 + em.a1 * em.sB) + kk * (d * d + Ds) * em.sA * em.sB;
}
}  // namespace

int ozaki_dgemm(cosma_amd_handle_t h, char ta, char tb, int64_t m, int64_t n, int64_t k, double alpha,
                const double *a, int64_t lda, const double *b, int64_t ldb, double beta, double *c,
                int64_t ldc, int method, double tol, int count, cosma_amd_ozaki_info *info, hipStream_t s,
                const OzakiFixedMemory *fixed) {
    if (info) *info = cosma_amd_ozaki_info{};
    if (m == 0 || n == 0) return COSMA_AMD_OK;
    if (k == 0 || alpha == 0.0) return launch_scale(m, n, beta, c, ldc, s);
    const bool strict = (method & COSMA_AMD_OZAKI_STRICT) != 0;
    method &= ~COSMA_AMD_OZAKI_STRICT;
    if (tol <= 0.0) tol = 1e-14;
    const int slices = method == COSMA_AMD_OZAKI_SLICES;
    const bool bf16 = method == COSMA_AMD_OZAKI_MODULI_BF16;   // residues as BF16, products on BF16 MFMA
    const int64_t esize = bf16 ? 2 : 1;
    const bool forced = count > 0;
    if (!slices && count > 0 && (count < 2 || count > MAX_MODULI)) {
        set_error("ozaki: %d moduli requested (2..%d supported)", count, MAX_MODULI);
        return COSMA_AMD_ERR_INVALID_ARG;
    }
    if (slices && count > I8_MAX_PAIRS) {
        set_error("ozaki: %d slices requested (1..%d supported)", count, I8_MAX_PAIRS);
        return COSMA_AMD_ERR_INVALID_ARG;
    }
    if (!slices && !bf16 && (double)k * 16384.0 >= 2147483648.0) {
        set_error("ozaki moduli: k = %lld too large for exact INT32 accumulation", (long long)k);
        return COSMA_AMD_ERR_UNSUPPORTED;
    }
    const int64_t Mp = round_up(m, 256), Np = round_up(n, 256), Kp = round_up(k, 256);
    const int64_t ldr = round_up(m, 16);                        // residue output ld
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    // Workspace layout (see steps 1 and 2 below): the statistics block `aux`,
    // then per pass the slices / residues of A and B and the product output.
    const int64_t nclear = 3 * (Mp + Np) + 2 * Kp + 1;
    const size_t norm_off = (size_t)nclear * 8 + (size_t)(Mp + Np) * 4;
    const size_t aux_bytes = al(norm_off + NORM_SLOTS * 8);
    const int64_t a_plane = Mp * Kp * esize + PLANE_PAD, b_plane = Np * Kp * esize + PLANE_PAD;
    const int64_t r_plane = ldr * n + PLANE_PAD;
    auto ws_bytes = [&](int cnt) {
        const size_t out_bytes = slices ? (size_t)ldr * n * 8 : (size_t)cnt * r_plane;
        return al((size_t)cnt * a_plane) + al((size_t)cnt * b_plane) + al(out_bytes);
    };
    const size_t span = (size_t)(2 * Kp + 2 * (Mp + Np) + 1) * 8 + (size_t)(Mp + Np) * 4;   // statistics read back
    // host staging: the statistics, then the NORM_SLOTS partial norms
    const size_t host_bytes = (span + 7) / 8 * 8 + NORM_SLOTS * 8;
    if (fixed) {
        // nothing is allocated: the largest count the call may use must fit
        const int most = count > 0 ? count : (slices ? I8_MAX_PAIRS : MAX_MODULI);
        if (aux_bytes + ws_bytes(most) > fixed->device_bytes || host_bytes > fixed->host_bytes) {
            set_error("ozaki (no allocation): m = %lld, n = %lld, k = %lld with up to %d %s need %zu bytes of "
                      "device scratch (%zu available) and %zu bytes of host staging (%zu available)",
                      (long long)m, (long long)n, (long long)k, most, slices ? "slices" : "moduli",
                      aux_bytes + ws_bytes(most), fixed->device_bytes, host_bytes, fixed->host_bytes);
            return COSMA_AMD_ERR_UNSUPPORTED;
        }
    }
    auto grow = [](DevBuf &buf, size_t need) -> int {
        if (buf.bytes >= need) return COSMA_AMD_OK;
        if (buf.ptr) CA_HIP(hipFree(buf.ptr));
        buf.ptr = nullptr;
        buf.bytes = 0;
        hipError_t e = dev_alloc(&buf.ptr, need);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            set_error("ozaki: workspace allocation of %zu bytes failed", need);
            return COSMA_AMD_ERR_ALLOC;
        }
        buf.bytes = need;
        return COSMA_AMD_OK;
    };
    int rc;

    // ---- 1. row / column maxima, sums of squares and 1-norms, per-k sums
    // of squares, exponents.  aux: maxA [Mp], maxB [Np], ksqA [Kp], ksqB [Kp],
    // rsqA [Mp], rsqB [Np], rabsA [Mp], rabsB [Np], flag [1] (cleared
    // together, 8 bytes each), then the exponents E [Mp] and F [Np], then
    // the NORM_SLOTS partial sums of the product's norm
    char *aux = nullptr;
    if (fixed) {
        aux = static_cast<char *>(fixed->device);
    } else {
        if ((rc = grow(h->ozaki_aux, aux_bytes))) return rc;
        aux = static_cast<char *>(h->ozaki_aux.ptr);
    }
    unsigned long long *maxA = reinterpret_cast<unsigned long long *>(aux);
    unsigned long long *maxB = maxA + Mp;
    double *ksqA = reinterpret_cast<double *>(maxB + Np);
    double *ksqB = ksqA + Kp;
    double *rsqA = ksqB + Kp;
    double *rsqB = rsqA + Mp;
    double *rabsA = rsqB + Np;
    double *rabsB = rabsA + Mp;
    int *flag = reinterpret_cast<int *>(rabsB + Np);
    int *E = reinterpret_cast<int *>(rabsB + Np + 1);
    int *F = E + Mp;
    double *nrm = reinterpret_cast<double *>(aux + norm_off);
    hipLaunchKernelGGL(clear_u64, dim3(grid_for(nclear)), dim3(256), 0, s, maxA, nclear);
    const bool a_rows_contig = ta == 'T', b_rows_contig = tb == 'N';   // op rows contiguous in k
    // grid: x over the contiguous dimension in 64s, y over the other in STATS_SPAN
    auto stats_grid = [](int64_t R, int64_t K, bool contig) {
        const int64_t fast = contig ? K : R, slow = contig ? R : K, span = 64 * (contig ? STATS_KC : STATS_RCH);
        return dim3((unsigned)((fast + span - 1) / span), (unsigned)((slow + STATS_SPAN - 1) / STATS_SPAN));
    };
    if (a_rows_contig) hipLaunchKernelGGL(stats_kernel<true>, stats_grid(m, k, true), dim3(256), 0, s, a, lda, m, k, maxA, ksqA, rsqA, rabsA);
    else hipLaunchKernelGGL(stats_kernel<false>, stats_grid(m, k, false), dim3(256), 0, s, a, lda, m, k, maxA, ksqA, rsqA, rabsA);
    if (b_rows_contig) hipLaunchKernelGGL(stats_kernel<true>, stats_grid(n, k, true), dim3(256), 0, s, b, ldb, n, k, maxB, ksqB, rsqB, rabsB);
    else hipLaunchKernelGGL(stats_kernel<false>, stats_grid(n, k, false), dim3(256), 0, s, b, ldb, n, k, maxB, ksqB, rsqB, rabsB);
    hipLaunchKernelGGL(exponents_kernel, dim3(grid_for(m)), dim3(256), 0, s, maxA, m, slices, E, flag);
    hipLaunchKernelGGL(exponents_kernel, dim3(grid_for(n)), dim3(256), 0, s, maxB, n, slices, F, flag);
    CA_HIP(hipGetLastError());
    // One read-back of the contiguous span ksqA .. F (each small copy costs
    // tens of microseconds of fixed overhead, which several copies paid in
    // turn while the GPU waited), into page-locked memory kept by the handle:
    // into pageable memory the runtime's staging added ~0.3 ms per call at
    // 8192^3.
    double *hb = nullptr;
    if (fixed) {
        hb = fixed->host;
    } else {
        if (h->ozaki_host.bytes < host_bytes) {
            if (h->ozaki_host.ptr) CA_HIP(hipHostFree(h->ozaki_host.ptr));
            h->ozaki_host.ptr = nullptr;
            h->ozaki_host.bytes = 0;
            const hipError_t e = pinned_alloc(&h->ozaki_host.ptr, host_bytes);
            if (e != hipSuccess) {
                (void)hipGetLastError();
                set_error("ozaki: page-locked statistics buffer of %zu bytes failed", host_bytes);
                return COSMA_AMD_ERR_ALLOC;
            }
            h->ozaki_host.bytes = host_bytes;
        }
        hb = static_cast<double *>(h->ozaki_host.ptr);
    }
    CA_HIP(hipMemcpyAsync(hb, ksqA, span, hipMemcpyDeviceToHost, s));
    CA_HIP(hipStreamSynchronize(s));
    const View<double> hkA{hb, k}, hkB{hb + Kp, k};
    const View<double> hrA{hb + 2 * Kp, m}, hrB{hb + 2 * Kp + Mp, n};
    const double *hl = hb + 2 * Kp + Mp + Np;
    const View<double> hlA{hl, m}, hlB{hl + Mp, n};
    int nonfinite;
    std::memcpy(&nonfinite, hl + Mp + Np, sizeof(int));
    const int *hEp = reinterpret_cast<const int *>(hl + Mp + Np + 1);
    const View<int> hE{hEp, m}, hF{hEp + Mp, n};
    if (nonfinite) {
        // Inf / NaN cannot be split; the FP64 MFMA DGEMM propagates them as IEEE does.
        if (info) {
            info->method = -1;
            info->error_bound = info->error_estimate = NAN;
            info->tolerance_met = -1;
        }
        return launch_dgemm(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, s);
    }

    // ---- precision.  The a-priori count: the error model's estimate
    // relative to the size of X that independent signs would give
    // (gamma * 2^-P); rounding every entry to P bits below its row / column
    // maximum gives gamma ~ 0.7 for entries of similar size, more when
    // entries sit far below their maxima.  After the products the norm of
    // the computed X replaces that assumed size (cancellation can make X far
    // smaller), and the count is raised and the product recomputed when the
    // error relative to it misses the tolerance: the estimate by default, the
    // rigorous bound with COSMA_AMD_OZAKI_STRICT.
    const ErrModel em = error_model(hE, hF, hkA, hkB, hlA, hlB, k);
    const double gamma = em.gamma;
    // Norm bounds of the scaled rows / columns for norm_bits: max over the
    // vectors of ||x|| 2^-E, slightly inflated for the rounding of the FP64
    // sums.  Only used when every exponent lies in [-480, 480], where the
    // squares of a vector's largest entries neither overflow nor underflow
    // (entries that do underflow are negligible beside them); otherwise P
    // comes from the entrywise bound alone.
    auto norm_max = [&](View<int> ex, View<double> rs) {
        const double *P2 = pow2_table();
        double mx = 0.0;
        for (int64_t i = 0; i < ex.size(); ++i) {
            if (ex[i] == ZERO_EXPONENT) continue;
            const double v = (ex[i] < -NORM_EXP_LIMIT || ex[i] > NORM_EXP_LIMIT) ? INFINITY
                                                                                  : scale2(P2, rs[i], -2 * ex[i]);
            mx = v > mx ? v : (v != v ? INFINITY : mx);
        }
        return std::sqrt(mx) * (1.0 + 1e-9);
    };
    const double sa = norm_max(hE, hrA), sb = norm_max(hF, hrB);
    const double hk = 0.5 * std::sqrt((double)k);
    auto bits = [&](int cnt) {
        const int pe = std::min(moduli_bits(cnt, k), MAX_P);
        return norm_bits(moduli_half_limit(cnt), hk, sa, sb, pe);
    };
    // The expected relative error (relative to d0): gamma 2^-P for moduli;
    // for slices the s-digit remainders and the dropped products of weight
    // 2^-7(s+2) and beyond give about gamma * 2 s 2^-7s, and measured errors
    // reached 1.2x that at 8192^3 (entries a = (u - 1/2) e^z), so twice that.
    // The a-priori count makes twice the expected error meet the tolerance;
    // after the products, 1.5 times the expected error relative to the
    // computed X' must meet it (the gap between the two factors keeps the
    // small difference between d0 and ||X'|| of operands without
    // cancellation from triggering a second pass).
    auto est = [&](int cnt) {
        return slices ? 2.0 * gamma * 2.0 * cnt * std::ldexp(1.0, -7 * cnt) : gamma * std::ldexp(1.0, -bits(cnt));
    };
    constexpr double PRIOR_FACTOR = 2.0, POST_FACTOR = 1.5;
    const int max_count = slices ? I8_MAX_PAIRS : MAX_MODULI;
    // relative errors after the products, with nx = ||X'||_F / 2^(ea + fb)
    auto post_est = [&](int cnt, double nx) { return em.zero ? 0.0 : est(cnt) * em.d0 / nx; };
    auto post_rig = [&](int cnt, double nx) {
        if (em.zero) return 0.0;
        const double r = rigorous_num(em, slices, slices ? cnt : bits(cnt));
        return nx > r ? r / (nx - r) : INFINITY;
    };
    auto meets = [&](int cnt, double nx) {
        return strict ? post_rig(cnt, nx) <= tol : POST_FACTOR * post_est(cnt, nx) <= tol;
    };
    int P = 0;
    if (!forced) {
        count = slices ? 1 : 2;
        if (!slices) {
            while (count < max_count && !(bits(count) >= 1 && PRIOR_FACTOR * est(count) <= tol)) ++count;
        } else {
            while (count < max_count && PRIOR_FACTOR * est(count) > tol) ++count;
        }
    }

    // the INT8 product kernel's duration (for the caller's roofline), timed
    // by two events around its launch (not with fixed memory, which creates
    // nothing)
    const bool timing = !fixed;
    if (timing)
        for (auto &ev : h->ozaki_ev)
            if (!ev) CA_HIP(hipEventCreate(&ev));
    const int shift = em.zero ? 0 : em.ea + em.fb + (int)std::ceil(std::log2((double)k + 1.0));
    int64_t products = 0;
    int passes = 0;
    double nx = 0.0, product_ms = 0.0;
    bool met = false;
    for (;;) {
        ++passes;
        if (!slices) {
            P = bits(count);
            if (P < 1) {
                set_error("ozaki: %d moduli leave no precision for k = %lld", count, (long long)k);
                return COSMA_AMD_ERR_UNSUPPORTED;
            }
        } else if ((double)count * (double)k * 16129.0 >= 2147483648.0) {
            // INT32 exactness of one weight: pairs * k * 127^2 < 2^31
            set_error("ozaki slices: k = %lld too large for exact INT32 accumulation", (long long)k);
            return COSMA_AMD_ERR_UNSUPPORTED;
        }

        // ---- workspace: slices / residues of A and B, then the INT8 GEMM output
        // Planes (one per modulus / slice) are PLANE_PAD bytes longer than their
        // data: at the usual power-of-two sizes the planes would otherwise start
        // at addresses 2^26 apart, and the split stores and recombination loads,
        // which touch the same offset of every plane together, would pile onto
        // the same memory channels.
        const size_t a_bytes = (size_t)count * a_plane, b_bytes = (size_t)count * b_plane;
        char *ws = nullptr;
        if (fixed) {
            ws = static_cast<char *>(fixed->device) + aux_bytes;
        } else {
            if ((rc = grow(h->ozaki_ws, ws_bytes(count)))) return rc;
            ws = static_cast<char *>(h->ozaki_ws.ptr);
        }
        int8_t *As = reinterpret_cast<int8_t *>(ws);
        int8_t *Bs = reinterpret_cast<int8_t *>(ws + al(a_bytes));
        void *out = ws + al(a_bytes) + al(b_bytes);

        // ---- 2. split
        auto split = [&](const double *x, int64_t ld, int64_t R, int64_t Rpad, bool contig, const int *ex,
                         int8_t *dst, int64_t stride) {
            SplitArgs sa{};
            sa.x = x; sa.ld = ld; sa.R = R; sa.K = k; sa.Kpad = Kp; sa.row_bytes = Kp * esize; sa.slice_stride = stride;
            sa.E = ex; sa.method = slices ? COSMA_AMD_OZAKI_SLICES : method; sa.esize = (int)esize;
            sa.count = count; sa.P = P; sa.out = dst;
            for (int l = 0; l < MAX_MODULI; ++l) { sa.m[l] = kModuli[l]; sa.inv_m[l] = 1.0 / kModuli[l]; }
            if (contig)
                hipLaunchKernelGGL(split_contig, dim3(grid_for(Rpad * (Kp / 16))), dim3(256), 0, s, sa, Rpad);
            else
                hipLaunchKernelGGL(split_strided, dim3((unsigned)(Rpad / 16), (unsigned)(Kp / 256)), dim3(256), 0, s, sa);
        };
        split(a, lda, m, Mp, a_rows_contig, E, As, a_plane);
        split(b, ldb, n, Np, b_rows_contig, F, Bs, b_plane);
        CA_HIP(hipGetLastError());

        // ---- 3. INT8 products
        I8Problem pr{};
        pr.M = m; pr.N = n; pr.K = Kp * esize;
        pr.a = As; pr.lda = Kp * esize; pr.a_stride = a_plane;
        pr.b = Bs; pr.ldb = Kp * esize; pr.b_stride = b_plane;
        pr.bf16 = bf16;
        pr.out = out; pr.ldo = ldr;
        // ---- 4. recombination with the norm of X (beta == 0: C written in
        // the same pass; otherwise C stays untouched until the check)
        hipLaunchKernelGGL(clear_u64, dim3(1), dim3(NORM_SLOTS), 0, s, reinterpret_cast<unsigned long long *>(nrm),
                           (int64_t)NORM_SLOTS);
        const int mode = beta == 0.0 ? EPI_WRITE_NORM : EPI_NORM;
        CrtConst cc;
        const dim3 gcrt((unsigned)((m + 256 * CRT_ROWS - 1) / (256 * CRT_ROWS)), (unsigned)std::min<int64_t>(n, 65535));
        auto recombine = [&](int md) {
            const int8_t *res = static_cast<const int8_t *>(out);
            if (crt_limbs(count) == 3) {
                if (md == EPI_WRITE) hipLaunchKernelGGL((crt_epilogue<3, EPI_WRITE>), gcrt, dim3(256), 0, s, cc, res, ldr, r_plane, m, n, P, E, F, alpha, beta, c, ldc, shift, nrm);
                else if (md == EPI_WRITE_NORM) hipLaunchKernelGGL((crt_epilogue<3, EPI_WRITE_NORM>), gcrt, dim3(256), 0, s, cc, res, ldr, r_plane, m, n, P, E, F, alpha, beta, c, ldc, shift, nrm);
                else hipLaunchKernelGGL((crt_epilogue<3, EPI_NORM>), gcrt, dim3(256), 0, s, cc, res, ldr, r_plane, m, n, P, E, F, alpha, beta, c, ldc, shift, nrm);
            } else {
                if (md == EPI_WRITE) hipLaunchKernelGGL((crt_epilogue<NLIMB, EPI_WRITE>), gcrt, dim3(256), 0, s, cc, res, ldr, r_plane, m, n, P, E, F, alpha, beta, c, ldc, shift, nrm);
                else if (md == EPI_WRITE_NORM) hipLaunchKernelGGL((crt_epilogue<NLIMB, EPI_WRITE_NORM>), gcrt, dim3(256), 0, s, cc, res, ldr, r_plane, m, n, P, E, F, alpha, beta, c, ldc, shift, nrm);
                else hipLaunchKernelGGL((crt_epilogue<NLIMB, EPI_NORM>), gcrt, dim3(256), 0, s, cc, res, ldr, r_plane, m, n, P, E, F, alpha, beta, c, ldc, shift, nrm);
            }
        };
        const dim3 gsl(grid_for(m * n));
        auto slices_out = [&](int md) {
            const double *acc = static_cast<const double *>(out);
            if (md == EPI_WRITE) hipLaunchKernelGGL((slices_epilogue<EPI_WRITE>), gsl, dim3(256), 0, s, acc, ldr, m, n, E, F, alpha, beta, c, ldc, shift, nrm);
            else if (md == EPI_WRITE_NORM) hipLaunchKernelGGL((slices_epilogue<EPI_WRITE_NORM>), gsl, dim3(256), 0, s, acc, ldr, m, n, E, F, alpha, beta, c, ldc, shift, nrm);
            else hipLaunchKernelGGL((slices_epilogue<EPI_NORM>), gsl, dim3(256), 0, s, acc, ldr, m, n, E, F, alpha, beta, c, ldc, shift, nrm);
        };
        int64_t pass_products = 0;
        if (!slices) {
            pr.out_stride = r_plane;
            I8Jobs jobs{};
            for (int l = 0; l < count; ++l) {
                jobs.npairs[l] = 1;
                jobs.pa[l][0] = (unsigned char)l;
                jobs.pb[l][0] = (unsigned char)l;
                jobs.modulus[l] = kModuli[l];
            }
            if (timing) CA_HIP(hipEventRecord(h->ozaki_ev[0], s));
            if ((rc = launch_i8gemm(pr, jobs, count, EPI_RES8, s))) return rc;
            if (timing) CA_HIP(hipEventRecord(h->ozaki_ev[1], s));
            pass_products = count;
            crt_constants(count, cc);
            recombine(mode);
        } else {
            pr.out_stride = 0;
            if (timing) CA_HIP(hipEventRecord(h->ozaki_ev[0], s));
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
                pass_products += np;
            }
            if (timing) CA_HIP(hipEventRecord(h->ozaki_ev[1], s));
            slices_out(mode);
        }
        CA_HIP(hipGetLastError());
        products += pass_products;
        double *parts = hb + (span + 7) / 8;   // NORM_SLOTS partial norms, after the statistics
        CA_HIP(hipMemcpyAsync(parts, nrm, NORM_SLOTS * sizeof(double), hipMemcpyDeviceToHost, s));
        CA_HIP(hipStreamSynchronize(s));
        if (timing) {
            float ms = 0.0f;
            CA_HIP(hipEventElapsedTime(&ms, h->ozaki_ev[0], h->ozaki_ev[1]));
            product_ms += ms;
        }
        double sum = 0.0;
        for (int q = 0; q < NORM_SLOTS; ++q) sum += parts[q];
        nx = std::sqrt(sum) * std::ldexp(1.0, shift - (em.ea + em.fb));   // ||X'||_F / 2^(ea + fb)
        met = meets(count, nx);
        const bool last = met || forced || count >= max_count || passes >= 3;
        if (last) {
            if (mode == EPI_NORM) {   // beta != 0: C is written now that the count is final
                if (!slices) recombine(EPI_WRITE);
                else slices_out(EPI_WRITE);
                CA_HIP(hipGetLastError());
            }
            break;
        }
        // smallest larger count that meets the tolerance relative to the
        // product just computed (the most precise one if none does)
        int next = count + 1;
        while (next < max_count && !meets(next, nx)) ++next;
        count = next;
    }
    if (info) {
        info->method = slices ? COSMA_AMD_OZAKI_SLICES : method;
        info->count = count;
        info->bits = slices ? 7 * count : P;
        info->int8_products = products;
        info->error_bound = post_rig(count, nx);
        info->error_estimate = post_est(count, nx);
        info->tolerance_met = met ? 1 : 0;
        info->passes = passes;
        info->product_ms = product_ms;
    }
    return COSMA_AMD_OK;
}

// ------------------------------------------------ no-allocation entry point
// Device scratch of the no-allocation split-precision call: the statistics
// and the residues / slices of one call, 64 MiB on every device the library
// runs on (held until the process ends).  Host code never uses the array's
// name as an address: each device's address comes from hipGetSymbolAddress
// (ozaki_scratch_address) and is what the kernels receive.
constexpr size_t OZAKI_SCRATCH_BYTES = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[OZAKI_SCRATCH_BYTES];

namespace {
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void *g_scratch_addr[MAX_DEVICES] = {};
// Host staging of the statistics read-back (shared by all no-allocation
// calls of the process, which it serialises).
constexpr size_t OZAKI_HOST_STAGING_BYTES = (size_t)1 << 20;
double g_host_staging[OZAKI_HOST_STAGING_BYTES / 8];
std::mutex g_host_staging_mu;
}  // namespace

int ozaki_scratch_address(int device, void **ptr) {
    *ptr = nullptr;
    if (device < 0 || device >= MAX_DEVICES) {
        set_error("ozaki scratch: device %d beyond the %d supported", device, MAX_DEVICES);
        return COSMA_AMD_ERR_UNSUPPORTED;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_addr[device]) {
        int cur = 0;
        CA_HIP(hipGetDevice(&cur));
        CA_HIP(hipSetDevice(device));
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        (void)hipSetDevice(cur);
        if (e != hipSuccess || !p) return hip_fail(e, "hipGetSymbolAddress(scratch)", __FILE__, __LINE__);
        g_scratch_addr[device] = p;
    }
    *ptr = g_scratch_addr[device];
    return COSMA_AMD_OK;
}

int ozaki_dgemm_noalloc(cosma_amd_handle_t h, char ta, char tb, int64_t m, int64_t n, int64_t k, double alpha,
                        const double *a, int64_t lda, const double *b, int64_t ldb, double beta, double *c,
                        int64_t ldc, int method, double tol, int count, cosma_amd_ozaki_info *info,
                        hipStream_t s) {
    if (!h->ozaki_scratch) {
        set_error("dgemm_ozaki_noalloc: the handle has no scratch address for device %d", h->device);
        return COSMA_AMD_ERR_UNSUPPORTED;
    }
    std::lock_guard<std::mutex> lock(g_host_staging_mu);
    const OzakiFixedMemory mem{h->ozaki_scratch, OZAKI_SCRATCH_BYTES, g_host_staging, OZAKI_HOST_STAGING_BYTES};
    return ozaki_dgemm(h, ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, method, tol, count, info, s, &mem);
}

}  // namespace cosma_amd
