// Grouped split-precision (moduli method) DGEMM plans: many independent
// device-resident problems C_p <- alpha_p op(A_p) op(B_p) + beta_p C_p, each
// FP64-accurate on the INT8 matrix cores, run as four grouped launches
// (exponents, split, INT8 products, recombination) that a plan can replay as
// one HIP graph.
//
// No counterpart in the reference, whose local multiplies are single plain
// GEMMs (src/cosma/local_multiply.cpp:299-381); this is the split-precision
// form of the FP64 grouped plans in grouped.hip.
//
// The method is the one of ozaki.hip (see there): per-row / per-column
// power-of-two exponents, integer inputs of P bits, symmetric residues modulo
// N pairwise-coprime moduli <= 256, exact INT32 products, Chinese-remainder
// recombination in exact FP64.  One moduli count N serves the whole plan (so
// the graph is fixed); each problem gets its own P from its k.  N is either
// given or chosen at plan creation from the tolerance and the operands present
// then: the per-problem error estimate gamma_p 2^-P_p of ozaki.hip (from the
// row / column maxima and the per-k sums of squares, computed on the device)
// must meet the tolerance, doubled, for every problem.  Later runs reuse N on
// whatever the operands hold; the exponents are recomputed on every run.
//
// Problems whose smallest dimension min(m, n, k) is below the plan's split
// threshold (default COSMA_AMD_OZAKI_SPLIT_MIN_DIM = 2048) are not split:
// the plan computes them with the FP64 MFMA grouped kernel of grouped.hip, in
// the same run (and graph), because below that size the FP64 kernel is both
// faster and FP64-accurate (its error bound is the FP64 GEMM's, k u |A||B|):
// the 4096-problem batch of dimensions 16..512 took 5.4 ms split at 14
// moduli against 0.88 ms in FP64.  A threshold of 0 splits every problem.
//
// Layout of the workspace (one allocation owned by the plan): per problem the
// residues of op(A) (N planes of Mp x Kp bytes, row r of op(A) k-contiguous),
// of op(B) (N planes of Np x Kp, column j k-contiguous), and of the products
// (N planes of ldr x n), where Mp, Np, Kp are m, n, k rounded up to 64 and
// ldr = m rounded up to 16; then the exponents E (Mp) and F (Np) and one
// non-finite flag per problem.  A problem whose operands hold Inf or NaN is
// computed by plain FP64 dot products in the recombination kernel, so IEEE
// propagation is kept.
#include <algorithm>
#include <cstring>
#include <cmath>
#include <numeric>
#include <vector>

#include "ozaki.hpp"

namespace cosma_amd {

struct OzProblem {
    int64_t m, n, k;
    double alpha, beta;
    const double *a;
    int64_t lda;
    const double *b;
    int64_t ldb;
    double *c;
    int64_t ldc;
    int ta, tb;              // 1: the stored operand is transposed
    int P;                   // bits of the integer inputs
    int scale_only;          // k == 0 or alpha == 0: C <- beta C
    int64_t Mp, Np, Kp;      // padded extents
    int8_t *ares, *bres;     // residues of op(A) rows / op(B) columns
    int8_t *ores;            // residues of the products
    int64_t oplane;          // bytes between product planes
    int64_t ldr;             // leading dimension of a product plane
    int *E, *F;              // exponents of the rows of op(A) / columns of op(B)
    int64_t tile_start;      // first INT8 tile of the problem in the product launch
};

static_assert(sizeof(OzProblem) % 8 == 0, "the int64 tile starts follow the descriptor array");

struct GzModuli {
    int count;
    int m[MAX_MODULI];
    double inv_m[MAX_MODULI];
    float inv_mf[MAX_MODULI];
};

}  // namespace cosma_amd

struct cosma_amd_grouped_ozaki_s {
    cosma_amd_handle_t h = nullptr;
    int nprob = 0;                         // problems computed by the split method
    cosma_amd_grouped_t fp64 = nullptr;    // the problems below the split threshold (FP64 grouped plan)
    int64_t nfp64 = 0;
    int moduli = 0;
    double estimate = 0.0;
    int64_t total_tiles = 0;
    int grid = 1;
    int vgroups = 1;   // max over the problems of max(Mp, Np) / 64
    cosma_amd::GzModuli md{};
    cosma_amd::CrtConst cc{};
    cosma_amd::OzProblem *dev = nullptr;   // descriptors, sorted by decreasing cost
    int *flags = nullptr;                  // one non-finite flag per problem
    unsigned long long *snorm = nullptr;   // two norm bounds per problem (bit patterns)
    int *emax = nullptr;                   // largest row / column exponent per problem (this run)
    double *xnorm = nullptr;               // sum of (X'_ij 2^-shift)^2 per problem (this run)
    std::vector<int64_t> original;         // caller's index of each sorted problem
    int *pbits = nullptr;                  // the bits P of every problem in the current run
    double limit = 0.0;                    // M / 2 of the plan's moduli, with margin
    int *expo = nullptr;                   // exponents E and F of every problem
    void *ws = nullptr;                    // residues
    hipStream_t capture = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace cosma_amd {
namespace {

constexpr int GZ_TILE = 64;   // INT8 product tile (rows x columns); k-step 64 bytes
constexpr int GZ_ROW = 80;    // LDS row stride in bytes: 16 consecutive rows start in 16 different bank groups

// Vector r of a column-major operand x seen as R vectors of length K: element
// (r, l) at x[l + r * ld] when `contig` (the vector runs along the stored
// column), else at x[r + l * ld].
__device__ __forceinline__ double vec_at(const double *x, int64_t ld, bool contig, int64_t r, int64_t l) {
    return contig ? x[l + r * ld] : x[r + l * ld];
}

// Exponents of the rows of op(A) and the columns of op(B), each problem's
// non-finite flag, and the norm bounds of its scaled vectors (gz_clear zeroes
// the flags and bounds first).  Workgroup (p, y): problem p, operand y & 1
// (0: op(A) rows, 1: op(B) columns), vectors [64 (y >> 1), 64 (y >> 1) + 64);
// entries in [R, Rp) get exponent 0.  A strided operand (vector r at
// x[r + l * ld]) is read by lane = vector, the four waves taking every fourth
// l with eight loads in flight; a contiguous one (x[l + r * ld]) by lanes
// along l, a wave taking 16 of the vectors four at a time, and reduced across
// the lanes.  snorm[2p + op] receives max ||x_r||^2 2^-2E_r over the vectors
// (as the bit pattern of a non-negative double; +inf when an exponent lies
// outside [-NORM_EXP_LIMIT, NORM_EXP_LIMIT], as in the single-problem path).
__global__ void gz_clear(int *flags, unsigned long long *snorm, int *emax, double *xnorm, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        flags[i] = 0;
        snorm[2 * i] = 0ull;
        snorm[2 * i + 1] = 0ull;
        emax[2 * i] = ZERO_EXPONENT;
        emax[2 * i + 1] = ZERO_EXPONENT;
        xnorm[i] = 0.0;
    }
}

// Besides the exponents and norm bounds, emax[2p + op] receives the largest
// exponent of the problem's rows of op(A) (op = 0) / columns of op(B) (1).
__global__ __launch_bounds__(256) void gz_exponents(const OzProblem *probs, int *flags, unsigned long long *snorm,
                                                    int *emax) {
    const OzProblem &q = probs[blockIdx.x];
    if (q.scale_only) return;
    const bool opA = (blockIdx.y & 1) == 0;
    const int64_t r0 = (int64_t)(blockIdx.y >> 1) * 64;
    const int64_t R = opA ? q.m : q.n, Rp = opA ? q.Mp : q.Np, K = q.k;
    if (r0 >= Rp) return;
    const double *x = opA ? q.a : q.b;
    const int64_t ld = opA ? q.lda : q.ldb;
    const bool contig = opA ? q.ta != 0 : q.tb == 0;
    __shared__ unsigned long long part[4][64];
    __shared__ double psq[4][64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    auto bits = [](double v) { return (unsigned long long)__double_as_longlong(fabs(v)); };
    if (!contig) {
        const int64_t r = r0 + lane;
        unsigned long long best = 0;
        double sq = 0.0;
        if (r < R) {
            int64_t l = wave;
            for (; l + 28 < K; l += 32) {
                double w[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) w[u] = x[r + (l + 4 * u) * ld];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const unsigned long long b = bits(w[u]);
                    best = b > best ? b : best;
                    sq = fma(w[u], w[u], sq);
                }
            }
            for (; l < K; l += 4) {
                const double w = x[r + l * ld];
                const unsigned long long b = bits(w);
                best = b > best ? b : best;
                sq = fma(w, w, sq);
            }
        }
        part[wave][lane] = best;
        psq[wave][lane] = sq;
    } else {
        for (int v0 = wave; v0 < 64; v0 += 16) {
            unsigned long long b[4] = {0, 0, 0, 0};
            double q2[4] = {0.0, 0.0, 0.0, 0.0};
            for (int64_t l = lane; l < K; l += 64) {
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int64_t r = r0 + v0 + 4 * u;
                    if (r < R) {
                        const double w = x[l + r * ld];
                        const unsigned long long c = bits(w);
                        b[u] = c > b[u] ? c : b[u];
                        q2[u] = fma(w, w, q2[u]);
                    }
                }
            }
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const unsigned long long o = __shfl_xor(b[u], off);
                    b[u] = o > b[u] ? o : b[u];
                    q2[u] += __shfl_xor(q2[u], off);
                }
            }
            if (lane < 4) {
                unsigned long long mine = b[0];
                double qs = q2[0];
#pragma unroll
                for (int u = 1; u < 4; ++u) {
                    mine = lane == u ? b[u] : mine;
                    qs = lane == u ? q2[u] : qs;
                }
                const int v = v0 + 4 * lane;
                part[v & 3][v >> 2] = mine;
                psq[v & 3][v >> 2] = qs;
            }
        }
    }
    __syncthreads();
    if (threadIdx.x < 64) {
        const int v = threadIdx.x;
        const int64_t r = r0 + v;
        double nrm = 0.0;   // ||x_r||^2 2^-2E_r of this vector
        int emx = ZERO_EXPONENT;
        if (r < Rp) {
            unsigned long long b;
            double sq;
            if (!contig) {
                b = part[0][v];
                sq = psq[0][v];
#pragma unroll
                for (int w = 1; w < 4; ++w) {
                    b = part[w][v] > b ? part[w][v] : b;
                    sq += psq[w][v];
                }
            } else {
                b = part[v & 3][v >> 2];
                sq = psq[v & 3][v >> 2];
            }
            if (r < R && b >= 0x7ff0000000000000ull) atomicOr(&flags[blockIdx.x], 1);   // inf or NaN
            const int e = r < R ? row_exponent(b, 0) : 0;
            (opA ? q.E : q.F)[r] = e;
            if (r < R && e != ZERO_EXPONENT)
                nrm = (e < -NORM_EXP_LIMIT || e > NORM_EXP_LIMIT) ? INFINITY : ldexp(sq, -2 * e);
            if (nrm != nrm) nrm = INFINITY;
            if (r < R && e > emx) emx = e;
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            nrm = fmax(nrm, __shfl_xor(nrm, off));
            emx = max(emx, __shfl_xor(emx, off));
        }
        if (v == 0) {
            atomicMax(&snorm[2 * blockIdx.x + (opA ? 0 : 1)], (unsigned long long)__double_as_longlong(nrm));
            atomicMax(&emax[2 * blockIdx.x + (opA ? 0 : 1)], emx);
        }
    }
}

// The bits P of every problem for this run: the entrywise bound's P (set at
// plan creation for the plan's moduli count) raised as far as the norm bound
// of the operands now present allows (norm_bits; the same rule and
// arithmetic as the single-problem path, whose results this keeps
// bit-identical).
__global__ void gz_bits(const OzProblem *probs, const int *flags, const unsigned long long *snorm, int *pbits,
                        int nprob, double limit) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nprob) return;
    const OzProblem &q = probs[i];
    int P = q.P;
    if (!q.scale_only && !flags[i]) {
        const double sa = sqrt(__longlong_as_double((long long)snorm[2 * i])) * (1.0 + 1e-9);
        const double sb = sqrt(__longlong_as_double((long long)snorm[2 * i + 1])) * (1.0 + 1e-9);
        P = norm_bits(limit, 0.5 * sqrt((double)q.k), sa, sb, q.P);
    }
    pbits[i] = P;
}

// Plan creation only: the error factor gamma_p of each problem (ozaki.hip's
// estimate: relative Frobenius error ~ gamma 2^-P), +inf when it cannot be
// formed, 0 for a zero product and for problems that do not use the moduli
// (scale-only, or flagged non-finite: those are computed in FP64).  Reads the
// exponents written by gz_exponents.
// d0[p] receives sqrt(sum_l ||op(A)(:, l)||^2 ||op(B)(l, :)||^2) 2^-(ea + fb),
// the size of X the estimate assumes (ea, fb the largest exponents).
__global__ __launch_bounds__(256) void gz_gamma(const OzProblem *probs, const int *flags, double *gamma,
                                                double *d0) {
    const OzProblem &q = probs[blockIdx.x];
    __shared__ int emax[2];
    __shared__ double part[5][256];
    const int t = threadIdx.x;
    if (q.scale_only || flags[blockIdx.x]) {
        if (t == 0) gamma[blockIdx.x] = d0[blockIdx.x] = 0.0;
        return;
    }
    if (t < 2) emax[t] = INT_MIN;
    __syncthreads();
    int ea = INT_MIN, fb = INT_MIN;
    for (int64_t i = t; i < q.m; i += blockDim.x) ea = max(ea, q.E[i]);
    for (int64_t j = t; j < q.n; j += blockDim.x) fb = max(fb, q.F[j]);
    atomicMax(&emax[0], ea);
    atomicMax(&emax[1], fb);
    __syncthreads();
    ea = emax[0];
    fb = emax[1];
    if (ea == ZERO_EXPONENT || fb == ZERO_EXPONENT) {
        if (t == 0) gamma[blockIdx.x] = d0[blockIdx.x] = 0.0;
        return;
    }
    double SA = 0.0, SB = 0.0, NA = 0.0, NB = 0.0, D = 0.0;
    for (int64_t i = t; i < q.m; i += blockDim.x) SA += ldexp(1.0, 2 * (q.E[i] - ea));
    for (int64_t j = t; j < q.n; j += blockDim.x) SB += ldexp(1.0, 2 * (q.F[j] - fb));
    for (int64_t l = t; l < q.k; l += blockDim.x) {
        double sa = 0.0, sb = 0.0;
        for (int64_t i = 0; i < q.m; ++i) {
            const double v = ldexp(vec_at(q.a, q.lda, q.ta != 0, i, l), -ea);
            sa = fma(v, v, sa);
        }
        for (int64_t j = 0; j < q.n; ++j) {
            const double v = ldexp(vec_at(q.b, q.ldb, q.tb == 0, j, l), -fb);
            sb = fma(v, v, sb);
        }
        NA += sa;
        NB += sb;
        D = fma(sa, sb, D);
    }
    part[0][t] = SA;
    part[1][t] = SB;
    part[2][t] = NA;
    part[3][t] = NB;
    part[4][t] = D;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w)
            for (int x = 0; x < 5; ++x) part[x][t] += part[x][t + w];
        __syncthreads();
    }
    if (t == 0) {
        const double d = part[4][0];
        const double g = sqrt((part[0][0] * part[3][0] + part[2][0] * part[1][0]) / (12.0 * d));
        gamma[blockIdx.x] = (d > 0.0 && isfinite(g)) ? g : INFINITY;
        d0[blockIdx.x] = sqrt(d);
    }
}

// Residues of the rows of op(A) and the columns of op(B), padding rows and k
// with zeros; workgroup (p, y) as in gz_exponents (problem p, operand y & 1,
// vectors [64 (y >> 1), 64 (y >> 1) + 64)).  Contiguous operand: a thread
// takes 16 consecutive k of one vector and consecutive threads continue
// along the vector, so the residue stores are contiguous runs.  Strided
// operand: 64 x 64 (vector x k) tiles are read along the vectors into LDS and
// transposed, four threads per vector writing 64 contiguous bytes per plane.
__global__ __launch_bounds__(256) void gz_split(const OzProblem *probs, const int *pbits, GzModuli md) {
    const OzProblem &q = probs[blockIdx.x];
    if (q.scale_only) return;
    const int P = pbits[blockIdx.x];
    const bool opA = (blockIdx.y & 1) == 0;
    const int64_t r0 = (int64_t)(blockIdx.y >> 1) * 64;
    const int64_t R = opA ? q.m : q.n, Rp = opA ? q.Mp : q.Np;
    if (r0 >= Rp) return;
    const double *x = opA ? q.a : q.b;
    const int64_t ld = opA ? q.lda : q.ldb;
    const bool contig = opA ? q.ta != 0 : q.tb == 0;
    const int *ex = opA ? q.E : q.F;
    int8_t *dst = opA ? q.ares : q.bres;
    const int64_t plane = Rp * q.Kp, nkc = q.Kp / 16;
    auto zeros = [&](int64_t r, int64_t k0) {
        for (int l = 0; l < md.count; ++l)
            *reinterpret_cast<i32x4 *>(dst + r * q.Kp + k0 + (int64_t)l * plane) = i32x4{0, 0, 0, 0};
    };
    if (contig) {
        const int64_t chunks = 64 * nkc;
        for (int64_t c = threadIdx.x; c < chunks; c += blockDim.x) {
            const int64_t r = r0 + c / nkc, k0 = 16 * (c % nkc);
            if (r >= R || k0 >= q.k) {   // padding: residues of zero
                zeros(r, k0);
                continue;
            }
            double v[16];
#pragma unroll
            for (int j = 0; j < 16; ++j) v[j] = k0 + j < q.k ? x[k0 + j + r * ld] : 0.0;
            moduli_residues16(v, ex[r], P, md.count, md.m, md.inv_m, dst + r * q.Kp + k0, plane);
        }
        return;
    }
    __shared__ double tile[64][65];
    const int t = threadIdx.x, tx = t & 63, ty = t >> 6;
    const int row = t >> 2, kc = t & 3;
    const int64_t r = r0 + row;
    const int e = ex[r];
    for (int64_t k0 = 0; k0 < q.Kp; k0 += 64) {
        if (k0 >= q.k || r0 >= R) {      // whole tile of padding (uniform branch)
            zeros(r, k0 + 16 * kc);
            continue;
        }
        __syncthreads();   // the previous tile's reads are done
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const int kk = ty + 4 * u;
            const int64_t rr = r0 + tx, k = k0 + kk;
            tile[kk][tx] = (rr < R && k < q.k) ? x[rr + k * ld] : 0.0;
        }
        __syncthreads();
        double v[16];
#pragma unroll
        for (int j = 0; j < 16; ++j) v[j] = tile[16 * kc + j][row];
        moduli_residues16(v, r < R ? e : 0, P, md.count, md.m, md.inv_m, dst + r * q.Kp + k0 + 16 * kc, plane);
    }
}

// The INT8 products of every problem and modulus: persistent workgroups walk
// the 64 x 64 output tiles (problems sorted by decreasing cost) and compute
// every modulus of a tile in turn, 2 x 2 waves of 32 x 32 (2 x 2
// v_mfma_i32_16x16x64_i8 each), k-step 64 double-buffered through LDS, exact
// INT32 accumulation, reduced modulo the modulus into symmetric int8
// residues.  The k-steps of all moduli of a tile form one pipeline, so the
// first step of modulus l + 1 is loaded while modulus l finishes and writes
// its residues, and the tile's problem is found once for all its moduli.
__global__ __launch_bounds__(256) void gz_products(const OzProblem *probs, int nprob, int64_t total, GzModuli md) {
    __shared__ __attribute__((aligned(16))) int8_t sa[2][GZ_TILE * GZ_ROW];
    __shared__ __attribute__((aligned(16))) int8_t sb[2][GZ_TILE * GZ_ROW];
    __shared__ __attribute__((aligned(16))) uint8_t stg[GZ_TILE * GZ_ROW];   // residue staging
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int wm = wave & 1, wn = wave >> 1, g = lane >> 4, li = lane & 15;
    const int lrow = t >> 2, lchunk = t & 3;   // this thread's 16 bytes of a 64 x 64-byte step
    // the descriptors are followed by their int64 tile_start values: the
    // search walks that compact array
    const int64_t *starts = reinterpret_cast<const int64_t *>(probs + nprob);
    for (int64_t v = blockIdx.x; v < total; v += gridDim.x) {
        // XCD-local chunks of 8 consecutive tiles (one problem, sharing
        // residue panels) on the XCD of workgroup v, as in the FP64 grouped
        // kernels; bijective on [0, total)
        constexpr int64_t NX = 8, CH = 8, span = NX * CH;
        const int64_t item =
            v >= total / span * span ? v : ((v / NX / CH) * NX + v % NX) * CH + (v / NX) % CH;
        int lo = 0, hi = nprob - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) / 2;
            if (starts[mid] <= item) lo = mid;
            else hi = mid - 1;
        }
        const OzProblem &q = probs[lo];
        const int64_t local = item - q.tile_start;
        const int64_t tm = q.Mp / GZ_TILE;
        const int64_t m0 = (local % tm) * GZ_TILE, n0 = (local / tm) * GZ_TILE;
        const int8_t *A0 = q.ares + (m0 + lrow) * q.Kp + 16 * lchunk;
        const int8_t *B0 = q.bres + (n0 + lrow) * q.Kp + 16 * lchunk;
        const int64_t aplane = q.Mp * q.Kp, bplane = q.Np * q.Kp;
        const int nsteps = (int)(q.Kp / 64), total_steps = md.count * nsteps;
        auto load = [&](int f, i32x4 &xa, i32x4 &xb) {
            const int l = f / nsteps, s = f - l * nsteps;
            xa = *reinterpret_cast<const i32x4 *>(A0 + (int64_t)l * aplane + 64 * s);
            xb = *reinterpret_cast<const i32x4 *>(B0 + (int64_t)l * bplane + 64 * s);
        };
        i32x4 acc[2][2];
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) acc[i][j] = i32x4{0, 0, 0, 0};
        __syncthreads();   // the previous tile's fragment reads are done
        i32x4 ra, rb;
        load(0, ra, rb);
        for (int f = 0; f < total_steps; ++f) {
            const int st = f & 1;
            *reinterpret_cast<i32x4 *>(sa[st] + lrow * GZ_ROW + 16 * lchunk) = ra;
            *reinterpret_cast<i32x4 *>(sb[st] + lrow * GZ_ROW + 16 * lchunk) = rb;
            __syncthreads();
            if (f + 1 < total_steps) load(f + 1, ra, rb);
            i32x4 fa[2], fb[2];
#pragma unroll
            for (int i = 0; i < 2; ++i)
                fa[i] = *reinterpret_cast<const i32x4 *>(sa[st] + (wm * 32 + 16 * i + li) * GZ_ROW + 16 * g);
#pragma unroll
            for (int j = 0; j < 2; ++j)
                fb[j] = *reinterpret_cast<const i32x4 *>(sb[st] + (wn * 32 + 16 * j + li) * GZ_ROW + 16 * g);
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(fa[i], fb[j], acc[i][j], 0, 0, 0);
            if ((f + 1) % nsteps) continue;
            // modulus l done.  Lane 16 g + li holds rows 4g .. 4g + 3 of
            // column li of each 16 x 16 block.  The 64 x 64 residues go
            // through the staging tile (column c at bytes [80 c, 80 c + 64):
            // the 4-byte writes of a wave hit 64 different banks) and leave
            // as 16-byte pieces, four per column.  The staging writes follow
            // the previous modulus's staging reads by at least one k-step
            // barrier.
            const int l = f / nsteps;
            const int mod = md.m[l];
            const float inv = md.inv_mf[l];
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j) {
                    const int rt = wm * 32 + 16 * i + 4 * g, ct = wn * 32 + 16 * j + li;
                    uint32_t w = 0;
#pragma unroll
                    for (int e = 0; e < 4; ++e)
                        w |= (uint32_t)(uint8_t)(int8_t)sym_mod(acc[i][j][e], mod, inv) << (8 * e);
                    *reinterpret_cast<uint32_t *>(stg + ct * GZ_ROW + rt) = w;
                    acc[i][j] = i32x4{0, 0, 0, 0};
                }
            __syncthreads();
            const int ct = t >> 2, c = t & 3;
            const int64_t col = n0 + ct, row = m0 + 16 * c;
            if (col < q.n && row < q.m) {
                const i32x4 val = *reinterpret_cast<const i32x4 *>(stg + ct * GZ_ROW + 16 * c);
                int8_t *o = q.ores + (int64_t)l * q.oplane + row + col * q.ldr;
                if (row + 16 <= q.m) {
                    *reinterpret_cast<i32x4 *>(o) = val;
                } else {
                    union { i32x4 q; int8_t b[16]; } u;
                    u.q = val;
                    for (int e = 0; e < 16 && row + e < q.m; ++e) o[e] = u.b[e];
                }
            }
        }
    }
}

// Recombination and epilogue: workgroup row blockIdx.x = problem, a thread
// takes 4 consecutive rows of one column.  Problems with k == 0 or alpha == 0
// only scale C; problems flagged non-finite are computed by FP64 dot products.
// xnorm[p] receives the sum of (X'_ij 2^-shift)^2 over the problem's
// recombined products X' (before alpha and beta), shift = ea + fb +
// ceil(log2(k + 1)) from emax, for the error check after a run.
template <int NL>
__global__ __launch_bounds__(256) void gz_recombine(const OzProblem *probs, const int *flags, const int *pbits,
                                                    CrtConst cc, const int *emax, double *xnorm) {
    const OzProblem &q = probs[blockIdx.x];
    const int P = pbits[blockIdx.x];
    const int64_t quads = (q.m + 3) / 4, items = quads * q.n;
    const bool plain = flags[blockIdx.x] != 0;
    const int shift = emax[2 * blockIdx.x] + emax[2 * blockIdx.x + 1] + (int)ceil(log2((double)q.k + 1.0));
    double ss = 0.0;
    for (int64_t it = threadIdx.x + (int64_t)blockIdx.y * blockDim.x; it < items;
         it += (int64_t)blockDim.x * gridDim.y) {
        const int64_t j = it / quads, i0 = (it - j * quads) * 4;
        double val[4] = {0.0, 0.0, 0.0, 0.0};
        if (q.scale_only) {
            // C <- beta C below
        } else if (plain) {
            for (int e = 0; e < 4 && i0 + e < q.m; ++e) {
                double acc = 0.0;
                for (int64_t l = 0; l < q.k; ++l)
                    acc = fma(vec_at(q.a, q.lda, q.ta != 0, i0 + e, l), vec_at(q.b, q.ldb, q.tb == 0, j, l), acc);
                val[e] = q.alpha * acc;
            }
        } else {
            double s[4][NL];
#pragma unroll
            for (int e = 0; e < 4; ++e)
#pragma unroll
                for (int x = 0; x < NL; ++x) s[e][x] = 0.0;
            const int8_t *rp = q.ores + i0 + j * q.ldr;
#pragma unroll 4
            for (int l = 0; l < cc.n; ++l) {
                const uint32_t word = *reinterpret_cast<const uint32_t *>(rp + (int64_t)l * q.oplane);
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const double c = (double)(int8_t)(word >> (8 * e));
#pragma unroll
                    for (int x = 0; x < NL; ++x) s[e][x] = fma(c, cc.w[l][x], s[e][x]);
                }
            }
            const int fj = q.F[j];
#pragma unroll
            for (int e = 0; e < 4; ++e)
                if (i0 + e < q.m) {
                    const double v = crt_value<NL>(s[e], cc);
                    const double x = ldexp(v, q.E[i0 + e] + fj - 2 * P - shift);
                    ss = fma(x, x, ss);
                    val[e] = q.alpha * ldexp(v, q.E[i0 + e] + fj - 2 * P);
                }
        }
        for (int e = 0; e < 4 && i0 + e < q.m; ++e) {
            double *cp = q.c + i0 + e + j * q.ldc;
            *cp = (q.beta == 0.0) ? val[e] : fma(q.beta, *cp, val[e]);
        }
    }
    // one atomic add per workgroup
    __shared__ double part[4];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) ss += __shfl_xor(ss, off);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = ss;
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(&xnorm[blockIdx.x], part[0] + part[1] + part[2] + part[3]);
}

constexpr int GZ_RECOMBINE_Y = 8;

void drop_graph(cosma_amd_grouped_ozaki_t p) {
    if (p->exec) (void)hipGraphExecDestroy(p->exec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    p->exec = nullptr;
    p->graph = nullptr;
}

bool on_device(const void *ptr) {
    hipPointerAttribute_t attr;
    if (hipPointerGetAttributes(&attr, ptr) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged;
}

inline int64_t up(int64_t x, int64_t a) { return (x + a - 1) / a * a; }

}  // namespace

int grouped_ozaki_destroy(cosma_amd_grouped_ozaki_t p) {
    if (!p) return COSMA_AMD_OK;
    (void)hipSetDevice(p->h->device);
    drop_graph(p);
    if (p->capture) (void)hipStreamDestroy(p->capture);
    if (p->fp64) grouped_destroy(p->fp64);
    if (p->dev) (void)hipFree(p->dev);
    if (p->flags) (void)hipFree(p->flags);
    if (p->snorm) (void)hipFree(p->snorm);
    if (p->pbits) (void)hipFree(p->pbits);
    if (p->emax) (void)hipFree(p->emax);
    if (p->xnorm) (void)hipFree(p->xnorm);
    if (p->expo) (void)hipFree(p->expo);
    if (p->ws) (void)hipFree(p->ws);
    delete p;
    return COSMA_AMD_OK;
}

int grouped_ozaki_run(cosma_amd_grouped_ozaki_t p, hipStream_t s) {
    CA_HIP(hipSetDevice(p->h->device));
    if (p->fp64) {
        const int rc = grouped_run(p->fp64, s);
        if (rc) return rc;
    }
    if (p->nprob == 0) return COSMA_AMD_OK;
    const dim3 gv((unsigned)p->nprob, (unsigned)(2 * p->vgroups));
    const dim3 g1((unsigned)((p->nprob + 255) / 256));
    hipLaunchKernelGGL(gz_clear, g1, dim3(256), 0, s, p->flags, p->snorm, p->emax, p->xnorm, p->nprob);
    hipLaunchKernelGGL(gz_exponents, gv, dim3(256), 0, s, p->dev, p->flags, p->snorm, p->emax);
    hipLaunchKernelGGL(gz_bits, g1, dim3(256), 0, s, p->dev, p->flags, p->snorm, p->pbits, p->nprob, p->limit);
    hipLaunchKernelGGL(gz_split, gv, dim3(256), 0, s, p->dev, p->pbits, p->md);
    if (p->total_tiles > 0)
        hipLaunchKernelGGL(gz_products, dim3((unsigned)p->grid), dim3(256), 0, s, p->dev, p->nprob, p->total_tiles,
                           p->md);
    const dim3 gr((unsigned)p->nprob, GZ_RECOMBINE_Y);
    if (crt_limbs(p->moduli) == 3)
        hipLaunchKernelGGL(gz_recombine<3>, gr, dim3(256), 0, s, p->dev, p->flags, p->pbits, p->cc, p->emax,
                           p->xnorm);
    else
        hipLaunchKernelGGL(gz_recombine<NLIMB>, gr, dim3(256), 0, s, p->dev, p->flags, p->pbits, p->cc, p->emax,
                           p->xnorm);
    CA_HIP(hipGetLastError());
    return COSMA_AMD_OK;
}

int grouped_ozaki_graph_run(cosma_amd_grouped_ozaki_t p, hipStream_t s) {
    if (p->nprob == 0 && !p->fp64) return COSMA_AMD_OK;
    CA_HIP(hipSetDevice(p->h->device));
    if (!p->exec) {
        if (!p->capture) CA_HIP(hipStreamCreateWithFlags(&p->capture, hipStreamNonBlocking));
        CA_HIP(hipStreamBeginCapture(p->capture, hipStreamCaptureModeThreadLocal));
        const int rc = grouped_ozaki_run(p, p->capture);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(p->capture, &g);
        if (rc != COSMA_AMD_OK) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        CA_HIP(e);
        p->graph = g;
        CA_HIP(hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0));
    }
    CA_HIP(hipGraphLaunch(p->exec, s));
    return COSMA_AMD_OK;
}

int grouped_ozaki_tune(cosma_amd_grouped_ozaki_t p, hipStream_t s) {
    CA_HIP(hipSetDevice(p->h->device));
    if (p->fp64) {
        const int rc = grouped_tune(p->fp64, s);
        drop_graph(p);
        if (rc) return rc;
    }
    if (p->nprob == 0 || p->total_tiles == 0) return COSMA_AMD_OK;
    hipEvent_t a, b;
    CA_HIP(hipEventCreate(&a));
    CA_HIP(hipEventCreate(&b));
    float best = -1.0f;
    int best_grid = p->grid;
    int rc = COSMA_AMD_OK;
    const int64_t cu = p->h->n_cu;
    for (int64_t g : {cu, 2 * cu, 4 * cu, 8 * cu, 16 * cu, p->total_tiles}) {
        p->grid = (int)std::max<int64_t>(1, std::min<int64_t>(g, p->total_tiles));
        if ((rc = grouped_ozaki_run(p, s))) break;   // warm-up
        if (hipEventRecord(a, s) != hipSuccess) { rc = COSMA_AMD_ERR_HIP; break; }
        for (int r = 0; r < 3 && rc == COSMA_AMD_OK; ++r) rc = grouped_ozaki_run(p, s);
        if (rc) break;
        if (hipEventRecord(b, s) != hipSuccess || hipEventSynchronize(b) != hipSuccess) {
            rc = COSMA_AMD_ERR_HIP;
            break;
        }
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, a, b);
        if (best < 0.0f || ms < best) {
            best = ms;
            best_grid = p->grid;
        }
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    p->grid = best_grid;
    drop_graph(p);
    if (rc) set_error("grouped ozaki: tuning failed");
    return rc;
}

int grouped_ozaki_create(cosma_amd_handle_t h, const cosma_amd_gemm_desc *d, int64_t count, int moduli,
                         double tol, int64_t split_min, cosma_amd_grouped_ozaki_t *out) {
    *out = nullptr;
    if (split_min < 0) {
        set_error("grouped ozaki: split threshold %lld is negative", (long long)split_min);
        return COSMA_AMD_ERR_INVALID_ARG;
    }
    if (count < 0 || count > 0x7fffffff || (count > 0 && !d)) {
        set_error("grouped ozaki: invalid problem list");
        return COSMA_AMD_ERR_INVALID_ARG;
    }
    if (moduli != 0 && (moduli < 2 || moduli > MAX_MODULI)) {
        set_error("grouped ozaki: %d moduli requested (2..%d supported)", moduli, MAX_MODULI);
        return COSMA_AMD_ERR_INVALID_ARG;
    }
    if (tol <= 0.0) tol = 1e-14;
    CA_HIP(hipSetDevice(h->device));
    std::vector<OzProblem> probs;
    std::vector<double> cost;
    std::vector<int64_t> orig;
    std::vector<cosma_amd_gemm_desc> plain;   // below the split threshold: FP64 grouped kernel
    for (int64_t i = 0; i < count; ++i) {
        const cosma_amd_gemm_desc &q = d[i];
        const char ta = norm_trans(q.transa), tb = norm_trans(q.transb);
        if (!ta || !tb || q.m < 0 || q.n < 0 || q.k < 0) {
            set_error("grouped ozaki: problem %lld has an invalid transpose or a negative dimension", (long long)i);
            return COSMA_AMD_ERR_INVALID_ARG;
        }
        if (q.m == 0 || q.n == 0) continue;
        const bool scale_only = (q.k == 0 || q.alpha == 0.0);
        const int64_t ar = ta == 'N' ? q.m : q.k, br = tb == 'N' ? q.k : q.n;
        if (q.ldc < std::max<int64_t>(1, q.m) || !q.c || !on_device(q.c) ||
            (!scale_only && (q.lda < std::max<int64_t>(1, ar) || q.ldb < std::max<int64_t>(1, br) || !q.a || !q.b ||
                             !on_device(q.a) || !on_device(q.b)))) {
            set_error("grouped ozaki: problem %lld has a leading dimension too small or an operand that is not "
                      "device memory", (long long)i);
            return COSMA_AMD_ERR_INVALID_ARG;
        }
        if (std::min(q.m, std::min(q.n, q.k)) < split_min) {
            plain.push_back(q);
            continue;
        }
        if (!scale_only && (double)q.k * 16384.0 >= 2147483648.0) {
            set_error("grouped ozaki: problem %lld: k = %lld too large for exact INT32 accumulation", (long long)i,
                      (long long)q.k);
            return COSMA_AMD_ERR_UNSUPPORTED;
        }
        OzProblem g{};
        g.m = q.m;
        g.n = q.n;
        g.k = q.k;
        g.alpha = q.alpha;
        g.beta = q.beta;
        g.a = q.a;
        g.lda = q.lda;
        g.b = q.b;
        g.ldb = q.ldb;
        g.c = q.c;
        g.ldc = q.ldc;
        g.ta = ta == 'T';
        g.tb = tb == 'T';
        g.scale_only = scale_only;
        g.Mp = up(q.m, GZ_TILE);
        g.Np = up(q.n, GZ_TILE);
        g.Kp = scale_only ? 0 : up(q.k, 64);
        g.ldr = up(q.m, 16);
        probs.push_back(g);
        cost.push_back(scale_only ? 0.0 : (double)q.m * q.n * q.k);
        orig.push_back(i);
    }
    std::vector<int> order(probs.size());
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return cost[x] > cost[y]; });
    std::vector<OzProblem> sorted;
    sorted.reserve(probs.size());
    for (int i : order) sorted.push_back(probs[i]);

    auto p = new cosma_amd_grouped_ozaki_s();
    p->h = h;
    p->nprob = (int)sorted.size();
    p->grid = 1;
    for (int i : order) p->original.push_back(orig[i]);
    auto fail = [&](int rc) {
        grouped_ozaki_destroy(p);
        return rc;
    };
    if (!plain.empty()) {
        const int rc = grouped_create(h, plain.data(), (int64_t)plain.size(), 0, &p->fp64);
        if (rc) return fail(rc);
        p->nfp64 = (int64_t)plain.size();
    }
    if (p->nprob == 0) {
        p->moduli = moduli;   // 0: no problem is split, no count was chosen
        *out = p;
        return COSMA_AMD_OK;
    }
    // vector groups of 64 per operand: the y extent of the exponent and split grids
    int64_t vg = 1;
    for (const auto &g : sorted) vg = std::max(vg, std::max(g.Mp, g.Np) / 64);
    if (2 * vg > 65535) {
        set_error("grouped ozaki: a problem with %lld rows or columns exceeds the launch limit",
                  (long long)(64 * vg));
        return fail(COSMA_AMD_ERR_UNSUPPORTED);
    }
    p->vgroups = (int)vg;
    auto alloc = [&](void **ptr, size_t bytes) -> int {
        const hipError_t e = dev_alloc(ptr, std::max<size_t>(bytes, 256));
        if (e != hipSuccess) {
            (void)hipGetLastError();
            *ptr = nullptr;
            set_error("grouped ozaki: allocation of %zu bytes failed (%s)", bytes, hipGetErrorString(e));
            return COSMA_AMD_ERR_ALLOC;
        }
        return COSMA_AMD_OK;
    };
    int rc;
    // descriptors, then their int64 first tile indices (the product kernel's
    // search array)
    const size_t ndesc = sorted.size() * sizeof(OzProblem), nstart = sorted.size() * sizeof(int64_t);
    if ((rc = alloc(reinterpret_cast<void **>(&p->dev), ndesc + nstart))) return fail(rc);
    if ((rc = alloc(reinterpret_cast<void **>(&p->flags), sorted.size() * sizeof(int)))) return fail(rc);
    if ((rc = alloc(reinterpret_cast<void **>(&p->snorm), 2 * sorted.size() * sizeof(unsigned long long))))
        return fail(rc);
    if ((rc = alloc(reinterpret_cast<void **>(&p->pbits), sorted.size() * sizeof(int)))) return fail(rc);
    if ((rc = alloc(reinterpret_cast<void **>(&p->emax), 2 * sorted.size() * sizeof(int)))) return fail(rc);
    if ((rc = alloc(reinterpret_cast<void **>(&p->xnorm), sorted.size() * sizeof(double)))) return fail(rc);

    // exponents first (their space does not depend on the moduli count)
    int64_t nexp = 0;
    for (const auto &g : sorted) nexp += g.Mp + g.Np;
    if ((rc = alloc(reinterpret_cast<void **>(&p->expo), (size_t)nexp * sizeof(int)))) return fail(rc);
    {
        int64_t off = 0;
        for (auto &g : sorted) {
            g.E = p->expo + off;
            off += g.Mp;
            g.F = p->expo + off;
            off += g.Np;
        }
    }
    auto upload = [&]() -> int {
        std::vector<int64_t> starts;
        for (const auto &g : sorted) starts.push_back(g.tile_start);
        CA_HIP(hipMemcpy(p->dev, sorted.data(), ndesc, hipMemcpyHostToDevice));
        CA_HIP(hipMemcpy(reinterpret_cast<char *>(p->dev) + ndesc, starts.data(), nstart, hipMemcpyHostToDevice));
        return COSMA_AMD_OK;
    };

    // error factors of the operands present now: they choose the count when
    // none is given, and give the plan's estimate either way
    if ((rc = upload())) return fail(rc);
    std::vector<double> gamma(sorted.size());
    std::vector<unsigned long long> hnorm(2 * sorted.size());
    {
        double *dgamma = nullptr;
        if ((rc = alloc(reinterpret_cast<void **>(&dgamma), 2 * sorted.size() * sizeof(double)))) return fail(rc);
        const dim3 gv((unsigned)p->nprob, (unsigned)(2 * p->vgroups));
        hipLaunchKernelGGL(gz_clear, dim3((unsigned)((p->nprob + 255) / 256)), dim3(256), 0, 0, p->flags, p->snorm,
                           p->emax, p->xnorm, p->nprob);
        hipLaunchKernelGGL(gz_exponents, gv, dim3(256), 0, 0, p->dev, p->flags, p->snorm, p->emax);
        hipLaunchKernelGGL(gz_gamma, dim3((unsigned)p->nprob), dim3(256), 0, 0, p->dev, p->flags, dgamma,
                           dgamma + sorted.size());
        hipError_t e = hipGetLastError();
        if (e == hipSuccess) e = hipMemcpy(gamma.data(), dgamma, gamma.size() * sizeof(double), hipMemcpyDeviceToHost);
        if (e == hipSuccess)
            e = hipMemcpy(hnorm.data(), p->snorm, hnorm.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost);
        (void)hipFree(dgamma);
        if (e != hipSuccess) {
            set_error("grouped ozaki: error estimation failed (%s)", hipGetErrorString(e));
            return fail(COSMA_AMD_ERR_HIP);
        }
    }
    // the doubled estimate of the worst problem for n moduli, with the bits
    // each problem gets in a run on the operands present now (gz_bits' rule)
    auto worst = [&](int n) {
        double w = 0.0;
        const double limit = moduli_half_limit(n);
        for (size_t i = 0; i < sorted.size(); ++i) {
            if (sorted[i].scale_only || gamma[i] == 0.0) continue;
            int P = std::min(moduli_bits(n, sorted[i].k), MAX_P);
            if (P < 1) return (double)INFINITY;
            double na, nb;
            std::memcpy(&na, &hnorm[2 * i], sizeof(double));
            std::memcpy(&nb, &hnorm[2 * i + 1], sizeof(double));
            P = norm_bits(limit, 0.5 * std::sqrt((double)sorted[i].k), std::sqrt(na) * (1.0 + 1e-9),
                          std::sqrt(nb) * (1.0 + 1e-9), P);
            w = std::max(w, 2.0 * gamma[i] * std::ldexp(1.0, -P));
        }
        return w;
    };
    if (moduli == 0) {
        moduli = 2;
        while (moduli < MAX_MODULI && !(worst(moduli) <= tol)) ++moduli;
    }
    p->estimate = worst(moduli);
    p->moduli = moduli;
    p->limit = moduli_half_limit(moduli);
    for (auto &g : sorted) {
        if (g.scale_only) continue;
        g.P = std::min(moduli_bits(moduli, g.k), MAX_P);
        if (g.P < 1) {
            set_error("grouped ozaki: %d moduli leave no precision for k = %lld", moduli, (long long)g.k);
            return fail(COSMA_AMD_ERR_UNSUPPORTED);
        }
    }

    // residues: per problem A planes, B planes, product planes (16-byte aligned)
    size_t bytes = 0;
    std::vector<size_t> offs(sorted.size());
    for (size_t i = 0; i < sorted.size(); ++i) {
        const auto &g = sorted[i];
        offs[i] = bytes;
        if (g.scale_only) continue;
        bytes += (size_t)moduli * (g.Mp + g.Np) * g.Kp + (size_t)moduli * up(g.ldr * g.n, 16);
    }
    if ((rc = alloc(&p->ws, bytes))) return fail(rc);
    int64_t tiles = 0;
    for (size_t i = 0; i < sorted.size(); ++i) {
        auto &g = sorted[i];
        g.tile_start = tiles;
        if (g.scale_only) continue;
        int8_t *base = static_cast<int8_t *>(p->ws) + offs[i];
        g.ares = base;
        g.bres = base + (size_t)moduli * g.Mp * g.Kp;
        g.ores = g.bres + (size_t)moduli * g.Np * g.Kp;
        g.oplane = up(g.ldr * g.n, 16);
        tiles += (g.Mp / GZ_TILE) * (g.Np / GZ_TILE);   // each tile computes every modulus
    }
    p->total_tiles = tiles;
    if ((rc = upload())) return fail(rc);
    p->md.count = moduli;
    for (int l = 0; l < MAX_MODULI; ++l) {
        p->md.m[l] = kModuli[l];
        p->md.inv_m[l] = 1.0 / kModuli[l];
        p->md.inv_mf[l] = 1.0f / (float)kModuli[l];
    }
    crt_constants(moduli, p->cc);
    p->grid = (int)std::max<int64_t>(1, std::min<int64_t>(tiles, (int64_t)h->n_cu * 8));
    *out = p;
    return COSMA_AMD_OK;
}

int grouped_ozaki_check(cosma_amd_grouped_ozaki_t p, double tol, double *worst, int64_t *worst_problem,
                        int *moduli_needed, hipStream_t s) {
    if (tol <= 0.0) tol = 1e-14;
    *worst = 0.0;
    *worst_problem = -1;
    *moduli_needed = p->moduli;
    if (p->nprob == 0) return COSMA_AMD_OK;
    CA_HIP(hipSetDevice(p->h->device));
    const size_t n = (size_t)p->nprob;
    double *dg = nullptr;
    if (dev_alloc(reinterpret_cast<void **>(&dg), 2 * n * sizeof(double)) != hipSuccess) {
        (void)hipGetLastError();
        set_error("grouped ozaki: check: allocation failed");
        return COSMA_AMD_ERR_ALLOC;
    }
    std::vector<double> gamma(n), d0(n), xn(n);
    std::vector<int> pb(n), em(2 * n), fl(n);
    std::vector<unsigned long long> sn(2 * n);
    std::vector<OzProblem> desc(n);
    hipLaunchKernelGGL(gz_gamma, dim3((unsigned)n), dim3(256), 0, s, p->dev, p->flags, dg, dg + n);
    hipError_t e = hipGetLastError();
    auto get = [&](void *dst, const void *src, size_t bytes) {
        if (e == hipSuccess) e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s);
    };
    get(gamma.data(), dg, n * sizeof(double));
    get(d0.data(), dg + n, n * sizeof(double));
    get(xn.data(), p->xnorm, n * sizeof(double));
    get(pb.data(), p->pbits, n * sizeof(int));
    get(em.data(), p->emax, 2 * n * sizeof(int));
    get(fl.data(), p->flags, n * sizeof(int));
    get(sn.data(), p->snorm, 2 * n * sizeof(unsigned long long));
    get(desc.data(), p->dev, n * sizeof(OzProblem));
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    (void)hipFree(dg);
    if (e != hipSuccess) return hip_fail(e, "grouped ozaki check", __FILE__, __LINE__);
    // expected relative error of problem i with P bits, relative to its X'
    auto est = [&](size_t i, int P) {
        const double nx = std::sqrt(xn[i]) * std::ldexp(1.0, (int)std::ceil(std::log2((double)desc[i].k + 1.0)));
        return nx > 0.0 ? gamma[i] * std::ldexp(1.0, -P) * d0[i] / nx : (double)INFINITY;
    };
    auto uses_moduli = [&](size_t i) { return !desc[i].scale_only && !fl[i] && gamma[i] != 0.0; };
    for (size_t i = 0; i < n; ++i) {
        if (!uses_moduli(i)) continue;
        const double v = est(i, pb[i]);
        if (!(v <= *worst)) {
            *worst = v;
            *worst_problem = p->original[i];
        }
    }
    // the smallest count whose bits meet the tolerance with 1.5 x the estimate
    // for every problem (the single-problem check's factor)
    int need = 2;
    for (; need <= MAX_MODULI; ++need) {
        const double limit = moduli_half_limit(need);
        bool ok = true;
        for (size_t i = 0; i < n && ok; ++i) {
            if (!uses_moduli(i)) continue;
            int P = std::min(moduli_bits(need, desc[i].k), MAX_P);
            if (P < 1) { ok = false; break; }
            double na, nb;
            std::memcpy(&na, &sn[2 * i], sizeof(double));
            std::memcpy(&nb, &sn[2 * i + 1], sizeof(double));
            P = norm_bits(limit, 0.5 * std::sqrt((double)desc[i].k), std::sqrt(na) * (1.0 + 1e-9),
                          std::sqrt(nb) * (1.0 + 1e-9), P);
            ok = 1.5 * est(i, P) <= tol;
        }
        if (ok) break;
    }
    *moduli_needed = need;   // MAX_MODULI + 1: no count meets the tolerance
    return COSMA_AMD_OK;
}

void grouped_ozaki_info(cosma_amd_grouped_ozaki_t p, int *moduli, double *estimate, int64_t *tiles, int *grid) {
    if (moduli) *moduli = p->moduli;
    if (estimate) *estimate = p->estimate;
    if (tiles) *tiles = p->total_tiles;
    if (grid) *grid = p->grid;
}

void grouped_ozaki_routing(cosma_amd_grouped_ozaki_t p, int64_t *split, int64_t *fp64) {
    if (split) *split = p->nprob;
    if (fp64) *fp64 = p->nfp64;
}

int grouped_ozaki_set_grid(cosma_amd_grouped_ozaki_t p, int grid) {
    if (grid < 1) {
        set_error("grouped ozaki: grid %d must be positive", grid);
        return COSMA_AMD_ERR_INVALID_ARG;
    }
    p->grid = grid;
    drop_graph(p);
    return COSMA_AMD_OK;
}

}  // namespace cosma_amd
