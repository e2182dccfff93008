// FP64 DGEMM on the gfx950 FP64 matrix cores.
//
//   C <- alpha * op(A) * op(B) + beta * C      (column-major, 64-bit extents)
//
// Replaces the GEMM that COSMA's local_multiply reaches on one rank
// (src/cosma/local_multiply.cpp:299-381 -> Tiled-MM gpu::gemm at :245-260,
// or cosma::gemm -> cblas_dgemm at src/cosma/blas.cpp:24-48).
//
// Work decomposition
//   * A 256-thread workgroup (4 waves of 64 lanes) owns a BM x BN = 128 x 128
//     block of C; each wave owns a 64 x 64 quarter, i.e. 4 x 4 MFMA blocks of
//     16 x 16, held as 16 accumulators of 4 doubles per lane.
//   * K is walked in steps of BK = 16.  Both operand tiles are staged in LDS
//     in a "k-pair" layout  tile[kp][row] = (x[row][2kp], x[row][2kp+1])  so
//     that one 16-byte LDS read per lane feeds two v_mfma_f64_16x16x4f64
//     (lane group g = lane/16 supplies k = 2kp and 2kp+1 of the same pair).
//     The order in which k is summed is a permutation of 0..K-1; it only
//     changes rounding, not the mathematical result.
//   * LDS is double-buffered: the global loads of step t+1 are issued into
//     registers before the MFMAs of step t and written to the other LDS
//     stage afterwards, so one barrier per step suffices.
//   * Operands whose rows are contiguous (A with 'N', B with 'T') and
//     operands whose k is contiguous (A with 'T', B with 'N') use two
//     different global-load patterns that land in the same LDS layout; all
//     four transpose combinations share the MFMA loop.
//
// MFMA v_mfma_f64_16x16x4f64 operand layout (lane l, g = l/16, i = l%16),
// decoded on gfx950 hardware:
//   a = A[i][g], b = B[g][i], d[e] = D[g + 4e][i]  for e = 0..3.
// The output rows of one lane are interleaved (g, g+4, g+8, g+12).  Feeding
// lane i the A row perm(i) = 4*(i%4) + i/4 instead of row i makes output
// element e of lane (g, i) hold row perm(g + 4e) = 4g + e, i.e. each lane
// ends with 4 consecutive rows of one column of C, which the epilogue
// writes as two 16-byte stores.
#include "common.hpp"

namespace cosma_amd {
namespace {

constexpr int BM = 128;
constexpr int BN = 128;
constexpr int BK = 16;
constexpr int KP = BK / 2;   // k-pairs per step
constexpr int NT = 256;      // threads per workgroup
constexpr int GROUP_M = 8;   // tile-rows per rasterisation group

struct alignas(16) Smem {
    f64x2 a[2][KP][BM];
    f64x2 b[2][KP][BN];
};

// Loads of one operand tile into 4 f64x2 registers per thread.
//   RowContig: element (r, k) at x[r + k * ld]   (A 'N', B 'T')
//   otherwise: element (r, k) at x[k + r * ld]   (A 'T', B 'N')
// r0/k0: tile origin; R/K: operand extents (rows of op(X), shared dim).
template <bool RowContig>
__device__ __forceinline__ void load_tile(const double *__restrict__ x, int64_t ld,
                                          int64_t r0, int64_t k0, int64_t R, int64_t K,
                                          bool vec_ok, f64x2 (&v)[4]) {
    const int t = threadIdx.x;
    const bool full = vec_ok && (r0 + BM <= R) && (k0 + BK <= K);
    if (RowContig) {
        // thread -> rows r = r0 + 2*(t%64) (+1), k-pairs kp = t/64 + 4q, q = 0,1
        const int64_t r = r0 + 2 * (t & 63);
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const int64_t k = k0 + 2 * ((t >> 6) + 4 * q);
            if (full) {
                v[2 * q] = *reinterpret_cast<const f64x2 *>(x + r + k * ld);
                v[2 * q + 1] = *reinterpret_cast<const f64x2 *>(x + r + (k + 1) * ld);
            } else {
#pragma unroll
                for (int dk = 0; dk < 2; ++dk) {
                    const bool kin = (k + dk) < K;
                    v[2 * q + dk].x = (kin && r < R) ? x[r + (k + dk) * ld] : 0.0;
                    v[2 * q + dk].y = (kin && r + 1 < R) ? x[r + 1 + (k + dk) * ld] : 0.0;
                }
            }
        }
    } else {
        // thread -> k-pair kp = t%8, rows r = r0 + t/8 + 32q, q = 0..3
        const int64_t k = k0 + 2 * (t & 7);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int64_t r = r0 + (t >> 3) + 32 * q;
            if (full) {
                v[q] = *reinterpret_cast<const f64x2 *>(x + k + r * ld);
            } else {
                const bool rin = r < R;
                v[q].x = (rin && k < K) ? x[k + r * ld] : 0.0;
                v[q].y = (rin && k + 1 < K) ? x[k + 1 + r * ld] : 0.0;
            }
        }
    }
}

template <bool RowContig, int ROWS>
__device__ __forceinline__ void store_tile(f64x2 (*tile)[ROWS], const f64x2 (&v)[4]) {
    const int t = threadIdx.x;
    if (RowContig) {
        const int r = 2 * (t & 63);
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const int kp = (t >> 6) + 4 * q;
            f64x2 lo, hi;
            lo.x = v[2 * q].x;  lo.y = v[2 * q + 1].x;   // row r:   (k, k+1)
            hi.x = v[2 * q].y;  hi.y = v[2 * q + 1].y;   // row r+1: (k, k+1)
            tile[kp][r] = lo;
            tile[kp][r + 1] = hi;
        }
    } else {
        const int kp = t & 7;
#pragma unroll
        for (int q = 0; q < 4; ++q) tile[kp][(t >> 3) + 32 * q] = v[q];
    }
}

template <bool TA, bool TB>
__global__ __launch_bounds__(NT) void dgemm_mfma_kernel(
    int64_t M, int64_t N, int64_t K, double alpha,
    const double *__restrict__ A, int64_t lda,
    const double *__restrict__ B, int64_t ldb,
    double beta, double *__restrict__ C, int64_t ldc,
    int vec_a, int vec_b, int vec_c) {
    __shared__ Smem sm;

    // ---- tile rasterisation: groups of GROUP_M tile-rows walk along n
    const int64_t tiles_m = (M + BM - 1) / BM;
    const int64_t tiles_n = (N + BN - 1) / BN;
    const int64_t bid = blockIdx.x;
    const int64_t per_group = (int64_t)GROUP_M * tiles_n;
    const int64_t group = bid / per_group;
    const int64_t first_m = group * GROUP_M;
    const int64_t gsize = min((int64_t)GROUP_M, tiles_m - first_m);
    const int64_t in_group = bid - group * per_group;
    const int64_t tm = first_m + in_group % gsize;
    const int64_t tn = in_group / gsize;
    const int64_t m0 = tm * BM, n0 = tn * BN;

    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    const int g = lane >> 4, li = lane & 15;
    const int arow = 4 * (li & 3) + (li >> 2);   // row permutation, see header

    f64x4 acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = f64x4{0.0, 0.0, 0.0, 0.0};

    const int64_t nk = (K + BK - 1) / BK;
    f64x2 ra[4], rb[4];

    load_tile<!TA>(A, lda, m0, 0, M, K, vec_a, ra);
    load_tile<TB>(B, ldb, n0, 0, N, K, vec_b, rb);
    store_tile<!TA, BM>(sm.a[0], ra);
    store_tile<TB, BN>(sm.b[0], rb);
    __syncthreads();

    for (int64_t kt = 0; kt < nk; ++kt) {
        const int st = kt & 1;
        const bool more = kt + 1 < nk;
        if (more) {
            load_tile<!TA>(A, lda, m0, (kt + 1) * BK, M, K, vec_a, ra);
            load_tile<TB>(B, ldb, n0, (kt + 1) * BK, N, K, vec_b, rb);
        }
#pragma unroll
        for (int ks = 0; ks < BK / 8; ++ks) {
            const int kp = 4 * ks + g;
            f64x2 af[4], bf[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) af[i] = sm.a[st][kp][wm * 64 + 16 * i + arow];
#pragma unroll
            for (int j = 0; j < 4; ++j) bf[j] = sm.b[st][kp][wn * 64 + 16 * j + li];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[i].x, bf[j].x, acc[i][j], 0, 0, 0);
                    acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[i].y, bf[j].y, acc[i][j], 0, 0, 0);
                }
        }
        if (more) {
            store_tile<!TA, BM>(sm.a[st ^ 1], ra);
            store_tile<TB, BN>(sm.b[st ^ 1], rb);
        }
        __syncthreads();
    }

    // ---- epilogue: lane holds rows 4g..4g+3 (contiguous) of column li
    const bool beta0 = (beta == 0.0);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int64_t row = m0 + wm * 64 + 16 * i + 4 * g;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int64_t col = n0 + wn * 64 + 16 * j + li;
            if (col >= N) continue;
            double *cp = C + row + col * ldc;
            const f64x4 r = acc[i][j];
            if (vec_c && row + 3 < M) {
                f64x2 *cv = reinterpret_cast<f64x2 *>(cp);
                f64x2 lo = {alpha * r[0], alpha * r[1]};
                f64x2 hi = {alpha * r[2], alpha * r[3]};
                if (!beta0) {
                    const f64x2 clo = cv[0], chi = cv[1];
                    lo.x = fma(beta, clo.x, lo.x); lo.y = fma(beta, clo.y, lo.y);
                    hi.x = fma(beta, chi.x, hi.x); hi.y = fma(beta, chi.y, hi.y);
                }
                cv[0] = lo;
                cv[1] = hi;
            } else {
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    if (row + e < M) {
                        double val = alpha * r[e];
                        if (!beta0) val = fma(beta, cp[e], val);
                        cp[e] = val;
                    }
                }
            }
        }
    }
}

// C <- beta * C; beta == 0 writes zeros without reading C.
__global__ void scale_kernel(int64_t M, int64_t N, double beta, double *__restrict__ C, int64_t ldc) {
    const int64_t total = M * N;
    for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < total;
         idx += (int64_t)gridDim.x * blockDim.x) {
        const int64_t col = idx / M, row = idx - col * M;
        double *p = C + row + col * ldc;
        *p = (beta == 0.0) ? 0.0 : beta * *p;
    }
}

inline bool aligned16(const void *p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; }

}  // namespace

int launch_scale(int64_t m, int64_t n, double beta, double *c, int64_t ldc, hipStream_t s) {
    if (m == 0 || n == 0 || beta == 1.0) return COSMA_AMD_OK;
    const int64_t total = m * n;
    int64_t blocks = (total + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(scale_kernel, dim3((unsigned)blocks), dim3(256), 0, s, m, n, beta, c, ldc);
    CA_HIP(hipGetLastError());
    return COSMA_AMD_OK;
}

int launch_dgemm(char ta, char tb, int64_t m, int64_t n, int64_t k, double alpha,
                 const double *a, int64_t lda, const double *b, int64_t ldb,
                 double beta, double *c, int64_t ldc, hipStream_t s) {
    if (m == 0 || n == 0) return COSMA_AMD_OK;
    if (k == 0 || alpha == 0.0) return launch_scale(m, n, beta, c, ldc, s);
    const int64_t tiles = ((m + BM - 1) / BM) * ((n + BN - 1) / BN);
    if (tiles > 0x7fffffffLL) {
        set_error("dgemm: %lld output tiles exceed the grid limit", (long long)tiles);
        return COSMA_AMD_ERR_UNSUPPORTED;
    }
    const int vec_a = aligned16(a) && (lda % 2 == 0);
    const int vec_b = aligned16(b) && (ldb % 2 == 0);
    const int vec_c = aligned16(c) && (ldc % 2 == 0);
    const dim3 grid((unsigned)tiles), block(NT);
    const bool TA = ta == 'T', TB = tb == 'T';
    if (!TA && !TB)
        hipLaunchKernelGGL((dgemm_mfma_kernel<false, false>), grid, block, 0, s, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, vec_a, vec_b, vec_c);
    else if (TA && !TB)
        hipLaunchKernelGGL((dgemm_mfma_kernel<true, false>), grid, block, 0, s, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, vec_a, vec_b, vec_c);
    else if (!TA && TB)
        hipLaunchKernelGGL((dgemm_mfma_kernel<false, true>), grid, block, 0, s, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, vec_a, vec_b, vec_c);
    else
        hipLaunchKernelGGL((dgemm_mfma_kernel<true, true>), grid, block, 0, s, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, vec_a, vec_b, vec_c);
    CA_HIP(hipGetLastError());
    return COSMA_AMD_OK;
}

}  // namespace cosma_amd
