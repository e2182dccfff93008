// FP64 DGEMM on the gfx950 FP64 matrix cores.
//
//   C <- alpha * op(A) * op(B) + beta * C      (column-major, 64-bit extents)
//
// Replaces the GEMM that COSMA's local_multiply reaches on one rank
// (src/cosma/local_multiply.cpp:299-381 -> Tiled-MM gpu::gemm at :245-260,
// or cosma::gemm -> cblas_dgemm at src/cosma/blas.cpp:24-48).
//
// Work decomposition
//   * A 256-thread workgroup (4 waves of 64 lanes, 2 x 2) owns a BM x BN block
//     of C; each wave owns a (BM/2) x (BN/2) quarter made of FM x FN MFMA
//     blocks of 16 x 16, each held as 4 doubles per lane.
//   * K is walked in steps of BK = 16.  Both operand tiles are staged in LDS
//     in a "k-pair" layout  tile[kp][row] = (x[row][2kp], x[row][2kp+1])  so
//     that one 16-byte LDS read per lane feeds two v_mfma_f64_16x16x4f64
//     (lane group g = lane/16 supplies k = 2kp and 2kp+1 of the same pair).
//     Rows are stored XOR-swizzled, tile[kp][row ^ kp ^ (bit 4 of row)] (flips the low
//     three row bits), with no padding: the MFMA reads (16 lanes on 16
//     consecutive rows of one kp) and the k-contiguous stores (eight lanes
//     per row, one per kp) are meant to hit distinct bank groups.  Measured
//     with SQ_LDS_BANK_CONFLICT on gfx950: a probe of the fragment reads
//     alone found them conflict-free (a one-entry row pad, the classic fix,
//     costs 0.76 conflict cycles per read); the whole TN 16384^3 kernel
//     still counts 0.89 conflict cycles per LDS instruction, 1.4% of its
//     cycles.
//     The order in which k is summed is a permutation of 0..K-1; it only
//     changes rounding, not the mathematical result.
//   * LDS is double-buffered: the global loads of step t+1 are issued into
//     registers before the MFMAs of step t and written to the other LDS
//     stage afterwards, so one barrier per step suffices.
//   * Operands whose rows are contiguous (A with 'N', B with 'T') and
//     operands whose k is contiguous (A with 'T', B with 'N') use two
//     different global-load patterns that land in the same LDS layout; all
//     four transpose combinations share the MFMA loop.
//   * Tile order is XCD-aware: the hardware hands workgroup b to XCD b % 8,
//     so workgroup ids are remapped so that each XCD receives a contiguous
//     range of logical tiles, and logical tiles walk GROUP_M tile-rows before
//     moving right.  The ~32 tiles resident on one XCD then form a compact
//     block of C whose A and B panels are shared through that XCD's L2.
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

constexpr int BK = 16;
constexpr int KP = BK / 2;   // k-pairs per step
constexpr int NT = 256;      // threads per workgroup
constexpr int NXCD = 8;      // XCDs on MI355X

// LDS position of row `row` in k-pair plane kp: the row XOR (kp ^ bit 4 of
// the row).  The flip stays inside an aligned group of 8 rows, so the 16
// consecutive rows of a fragment read land on 16 different 16-byte bank
// groups; bit 4 separates the two halves of the 32 rows that 16 lanes store
// at once for a row-contiguous operand (rows 2l and 2l + 1 per lane), which
// with kp alone met in pairs (a 2-way conflict on every such store).
__device__ __forceinline__ int lds_swz(int row, int kp) { return row ^ kp ^ ((row >> 4) & 1); }

// Loads of one operand tile (ROWS x BK) into registers.
//   RowContig: element (r, k) at x[r + k * ld]   (A 'N', B 'T')
//     work item idx -> row pair rp = idx % (ROWS/2), k-pair kp = idx / (ROWS/2);
//     two loads (k = 2kp and 2kp+1), each covering rows 2rp, 2rp+1.
//   otherwise: element (r, k) at x[k + r * ld]   (A 'T', B 'N')
//     work item idx -> (row, k-pair) by kmap; one load.
// Rows at or beyond R are clamped to a valid row: they only feed rows (or
// columns) of C that the epilogue never stores, so their values are
// irrelevant and no branch is needed.  Only k must be exact: full steps
// (load) need no k check; the one partial step at the end of K (load_tail)
// zero-fills k >= K.
// VEC: 16-byte loads; requires 16-byte aligned x, even ld and, for
// RowContig, even R (so a row pair never straddles the end of a column).
template <bool RowContig, int ROWS, bool VEC>
struct TileIO {
    static constexpr int ITEMS = RowContig ? (ROWS / 2) * KP : ROWS * KP;
    static constexpr int PER_T = ITEMS / NT;           // work items per thread
    static constexpr int NREG = RowContig ? 2 * PER_T : PER_T;
    static_assert(ITEMS % NT == 0, "tile must divide evenly over the workgroup");
    static_assert(RowContig || ROWS % 16 == 0, "k-contiguous tiles need 16-row groups");

    // k-contiguous operands: work item idx -> (row, k-pair).  Eight lanes read
    // the 128 bytes of one row (k-pairs 0..7), and each group of 16 lanes
    // (one LDS store pass) takes rows r and r + 8, so the 16 stores land on
    // 16 different 16-byte bank groups of the padded [kp][row] tile
    // (address mod 256 = 16 * (kp + row) mod 256).  Wave-instruction pairs
    // (W even / odd) cover 16 consecutive rows.
    __device__ __forceinline__ static void kmap(int idx, int &row, int &kp) {
        const int lane = idx & 63, W = idx >> 6;
        kp = lane & 7;
        row = 16 * (W >> 1) + 4 * (W & 1) + (lane >> 4) + 8 * ((lane >> 3) & 1);
    }

    __device__ __forceinline__ static void load(const double *__restrict__ x, int64_t ld,
                                                int64_t r0, int64_t k0, int64_t R,
                                                f64x2 (&v)[NREG]) {
        const int t = threadIdx.x;
#pragma unroll
        for (int q = 0; q < PER_T; ++q) {
            const int idx = t + NT * q;
            if (RowContig) {
                const int64_t k = k0 + 2 * (idx / (ROWS / 2));
                if (VEC) {
                    const int64_t r = min(r0 + 2 * (idx % (ROWS / 2)), R - 2);
                    v[2 * q] = *reinterpret_cast<const f64x2 *>(x + r + k * ld);
                    v[2 * q + 1] = *reinterpret_cast<const f64x2 *>(x + r + (k + 1) * ld);
                } else {
                    const int64_t ra = min(r0 + 2 * (idx % (ROWS / 2)), R - 1);
                    const int64_t rb = min(r0 + 2 * (idx % (ROWS / 2)) + 1, R - 1);
                    v[2 * q].x = x[ra + k * ld];
                    v[2 * q].y = x[rb + k * ld];
                    v[2 * q + 1].x = x[ra + (k + 1) * ld];
                    v[2 * q + 1].y = x[rb + (k + 1) * ld];
                }
            } else {
                int row, kp;
                kmap(idx, row, kp);
                const int64_t k = k0 + 2 * kp;
                const int64_t r = min(r0 + row, R - 1);
                if (VEC) {
                    v[q] = *reinterpret_cast<const f64x2 *>(x + k + r * ld);
                } else {
                    v[q].x = x[k + r * ld];
                    v[q].y = x[k + 1 + r * ld];
                }
            }
        }
    }

    __device__ __forceinline__ static void load_tail(const double *__restrict__ x, int64_t ld,
                                                     int64_t r0, int64_t k0, int64_t R, int64_t K,
                                                     f64x2 (&v)[NREG]) {
        const int t = threadIdx.x;
#pragma unroll
        for (int q = 0; q < PER_T; ++q) {
            const int idx = t + NT * q;
            if (RowContig) {
                const int64_t ra = min(r0 + 2 * (idx % (ROWS / 2)), R - 1);
                const int64_t rb = min(r0 + 2 * (idx % (ROWS / 2)) + 1, R - 1);
                const int64_t k = k0 + 2 * (idx / (ROWS / 2));
#pragma unroll
                for (int dk = 0; dk < 2; ++dk) {
                    const bool kin = (k + dk) < K;
                    v[2 * q + dk].x = kin ? x[ra + (k + dk) * ld] : 0.0;
                    v[2 * q + dk].y = kin ? x[rb + (k + dk) * ld] : 0.0;
                }
            } else {
                int row, kp;
                kmap(idx, row, kp);
                const int64_t k = k0 + 2 * kp;
                const int64_t r = min(r0 + row, R - 1);
                v[q].x = (k < K) ? x[k + r * ld] : 0.0;
                v[q].y = (k + 1 < K) ? x[k + 1 + r * ld] : 0.0;
            }
        }
    }

    __device__ __forceinline__ static void store(f64x2 (*tile)[ROWS], const f64x2 (&v)[NREG]) {
        const int t = threadIdx.x;
#pragma unroll
        for (int q = 0; q < PER_T; ++q) {
            const int idx = t + NT * q;
            if (RowContig) {
                const int r = 2 * (idx % (ROWS / 2));
                const int kp = idx / (ROWS / 2);
                f64x2 lo, hi;
                lo.x = v[2 * q].x;  lo.y = v[2 * q + 1].x;   // row r:   (k, k+1)
                hi.x = v[2 * q].y;  hi.y = v[2 * q + 1].y;   // row r+1: (k, k+1)
                tile[kp][lds_swz(r, kp)] = lo;
                tile[kp][lds_swz(r + 1, kp)] = hi;
            } else {
                int row, kp;
                kmap(idx, row, kp);
                tile[kp][lds_swz(row, kp)] = v[q];
            }
        }
    }
};

// Maps a hardware workgroup id to a logical tile id so that XCD x (which
// receives ids b with b % 8 == x) owns the contiguous logical range
// [start_x, start_x + count_x).  Bijective on [0, T).
__device__ __forceinline__ int64_t xcd_remap(int64_t b, int64_t T) {
    const int64_t xcd = b % NXCD, local = b / NXCD;
    const int64_t q = T / NXCD, r = T % NXCD;
    return (xcd < r) ? xcd * (q + 1) + local : r * (q + 1) + (xcd - r) * q + local;
}

template <int BM, int BN>
struct TileSmem {
    f64x2 a[2][KP][BM];
    f64x2 b[2][KP][BN];
};

// One BM x BN tile of C at (m0, n0): the whole k loop and the epilogue,
// executed by the 256 threads of the workgroup.  Used by the single-GEMM
// kernel (one tile per workgroup) and by the grouped kernel (a persistent
// loop over the tiles of many problems); the caller provides the LDS tiles
// and must not touch them while this runs.
// EDGE: skip the MFMAs of 16 x 16 blocks that lie wholly outside C (tiles on
// the bottom / right edge of small problems); their waves leave the SIMD to
// the other workgroups resident on the CU.
template <int BM, int BN, bool TA, bool TB, bool VEC, bool EDGE = false>
__device__ __forceinline__ void gemm_tile(
    int64_t M, int64_t N, int64_t K, double alpha,
    const double *__restrict__ A, int64_t lda,
    const double *__restrict__ B, int64_t ldb,
    double beta, double *__restrict__ C, int64_t ldc,
    int vec_c, int64_t m0, int64_t n0, TileSmem<BM, BN> &sm) {
    constexpr int FM = BM / 32, FN = BN / 32;    // MFMA blocks per wave
    using IOA = TileIO<!TA, BM, VEC>;
    using IOB = TileIO<TB, BN, VEC>;
    auto &sa = sm.a;
    auto &sb = sm.b;

    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    const int g = lane >> 4, li = lane & 15;
    const int arow = 4 * (li & 3) + (li >> 2);   // row permutation, see header
    const int abase = wm * (BM / 2) + arow;
    const int bbase = wn * (BN / 2) + li;

    // The accumulators are zeroed by an MFMA (0 * 0 + 0) rather than by
    // VALU moves: values that originate from an MFMA are allocated to AGPRs,
    // which keeps them there across the k loop; zeroing with moves led the
    // compiler to carry them in VGPRs and copy all of them to and from AGPRs
    // in every iteration.
    f64x4 acc[FM][FN];
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j)
            acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(0.0, 0.0, f64x4{0.0, 0.0, 0.0, 0.0}, 0, 0, 0);

    // Fragment registers: F[h] holds the A/B fragments of k-sub-step h (two
    // MFMA k-slices each) so that LDS reads of one sub-step overlap the
    // MFMAs of the other.
    f64x2 af[2][FM], bf[2][FN];
    auto read_frags = [&](int st, int ks, int h) {
        const int kp = 4 * ks + g;
#pragma unroll
        for (int i = 0; i < FM; ++i) af[h][i] = sa[st][kp][lds_swz(abase + 16 * i, kp)];
#pragma unroll
        for (int j = 0; j < FN; ++j) bf[h][j] = sb[st][kp][lds_swz(bbase + 16 * j, kp)];
    };
    // blocks of this wave that hold rows / columns of C (all of them unless EDGE)
    const int64_t wrow0 = m0 + wm * (BM / 2), wcol0 = n0 + wn * (BN / 2);
    const int rows_in = EDGE ? (int)max((int64_t)0, min((int64_t)FM, (M - wrow0 + 15) / 16)) : FM;
    const int cols_in = EDGE ? (int)max((int64_t)0, min((int64_t)FN, (N - wcol0 + 15) / 16)) : FN;
    const bool whole = !EDGE || (rows_in == FM && cols_in == FN);
    auto mfma_frags = [&](int h) {
        if (whole) {
#pragma unroll
            for (int i = 0; i < FM; ++i)
#pragma unroll
                for (int j = 0; j < FN; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[h][i].x, bf[h][j].x, acc[i][j], 0, 0, 0);
#pragma unroll
            for (int i = 0; i < FM; ++i)
#pragma unroll
                for (int j = 0; j < FN; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[h][i].y, bf[h][j].y, acc[i][j], 0, 0, 0);
        } else {
#pragma unroll
            for (int i = 0; i < FM; ++i)
#pragma unroll
                for (int j = 0; j < FN; ++j)
                    if (i < rows_in && j < cols_in) {
                        acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[h][i].x, bf[h][j].x, acc[i][j], 0, 0, 0);
                        acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[h][i].y, bf[h][j].y, acc[i][j], 0, 0, 0);
                    }
        }
    };

    // LDS stores per thread and step, MFMAs per k-sub-step
    constexpr int N_WRITES = (TA ? 1 : 2) * IOA::PER_T + (TB ? 2 : 1) * IOB::PER_T;
    constexpr int MFMA_HALF = 2 * FM * FN;
    constexpr int MFMA_PER_WRITE = MFMA_HALF / (N_WRITES + 1);
    // global loads per thread and step (VEC: 16-byte loads; else 8-byte)
    constexpr int N_LOADS = (VEC ? 1 : 2) * (IOA::NREG + IOB::NREG);
    constexpr int MFMA_PER_LOAD = MFMA_HALF / (N_LOADS + 1) > 0 ? MFMA_HALF / (N_LOADS + 1) : 1;
    const int64_t nk_full = K / BK;          // full k-steps
    const bool has_tail = (K % BK) != 0;     // one partial step after them
    f64x2 ra[IOA::NREG], rb[IOB::NREG];
    auto load_full = [&](int64_t kn) {
        IOA::load(A, lda, m0, kn * BK, M, ra);
        IOB::load(B, ldb, n0, kn * BK, N, rb);
    };
    auto load_tail = [&](int64_t kn) {
        IOA::load_tail(A, lda, m0, kn * BK, M, K, ra);
        IOB::load_tail(B, ldb, n0, kn * BK, N, K, rb);
    };
    auto store_regs = [&](int st) {
        IOA::store(sa[st], ra);
        IOB::store(sb[st], rb);
    };

    // Pipeline over the full steps.  Entering iteration kt, LDS stage kt&1
    // holds step kt, F[0] holds its first k-sub-step and the registers hold
    // step kt+1.  Iteration kt: read sub-step 1 into F[1]; MFMAs of F[0];
    // store step kt+1 into the other stage (last read before the previous
    // barrier); load step kt+2; barrier; read sub-step 0 of step kt+1 into
    // F[0]; MFMAs of F[1].  Each batch of fragment reads is issued 64 MFMAs
    // before it is used, and each global load a full step before it is
    // stored.  The body has no branches (the loads of the last iterations
    // re-read the last full step; the extra store and fragment reads are
    // never used), which keeps the compiler's wait counts partial, and the
    // scheduling barriers stop the scheduler from sinking fragment reads to
    // their first use.
    if (nk_full > 0) {
        load_full(0);
        store_regs(0);
        load_full(min((int64_t)1, nk_full - 1));
        __syncthreads();
        read_frags(0, 0, 0);
        __builtin_amdgcn_sched_barrier(0);
        for (int64_t kt = 0; kt < nk_full; ++kt) {
            const int st = (int)(kt & 1);
            read_frags(st, 1, 1);
            __builtin_amdgcn_sched_barrier(0);
            mfma_frags(0);
            store_regs(st ^ 1);
            // spread the LDS stores over the MFMAs instead of stalling the
            // wave on them after the last MFMA
#pragma unroll
            for (int w = 0; w < N_WRITES; ++w) {
                __builtin_amdgcn_sched_group_barrier(0x008, MFMA_PER_WRITE, 0);   // MFMA
                __builtin_amdgcn_sched_group_barrier(0x200, 1, 0);                // DS write
            }
            __builtin_amdgcn_sched_group_barrier(0x008, MFMA_HALF - N_WRITES * MFMA_PER_WRITE, 0);
            __builtin_amdgcn_sched_barrier(0);
            __syncthreads();
            read_frags(st ^ 1, 0, 0);
            __builtin_amdgcn_sched_barrier(0);
            mfma_frags(1);
            load_full(min(kt + 2, nk_full - 1));
            // and the global loads over the MFMAs of the second sub-step
#pragma unroll
            for (int w = 0; w < N_LOADS; ++w) {
                __builtin_amdgcn_sched_group_barrier(0x008, MFMA_PER_LOAD, 0);    // MFMA
                __builtin_amdgcn_sched_group_barrier(0x020, 1, 0);                // VMEM read
            }
            __builtin_amdgcn_sched_group_barrier(0x008, MFMA_HALF - N_LOADS * MFMA_PER_LOAD, 0);
            __builtin_amdgcn_sched_barrier(0);
        }
    }
    if (has_tail) {
        // The partial last step, zero-filled beyond K, through stage
        // nk_full & 1 (whose last reads finish before this barrier).
        load_tail(nk_full);
        __syncthreads();
        const int st = (int)(nk_full & 1);
        store_regs(st);
        __syncthreads();
        read_frags(st, 0, 0);
        read_frags(st, 1, 1);
        mfma_frags(0);
        mfma_frags(1);
    }

    // ---- epilogue: lane holds rows 4g..4g+3 (contiguous) of column li
    const bool beta0 = (beta == 0.0);
#pragma unroll
    for (int i = 0; i < FM; ++i) {
        const int64_t row = m0 + wm * (BM / 2) + 16 * i + 4 * g;
#pragma unroll
        for (int j = 0; j < FN; ++j) {
            const int64_t col = n0 + wn * (BN / 2) + 16 * j + li;
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

template <int BM, int BN, bool TA, bool TB, bool VEC>
__global__ __launch_bounds__(NT) void dgemm_mfma_kernel(
    int64_t M, int64_t N, int64_t K, double alpha,
    const double *__restrict__ A, int64_t lda,
    const double *__restrict__ B, int64_t ldb,
    double beta, double *__restrict__ C, int64_t ldc,
    int vec_c, int group_m) {
    __shared__ TileSmem<BM, BN> sm;
    // ---- tile rasterisation (XCD-aware, grouped along m)
    const int64_t tiles_m = (M + BM - 1) / BM;
    const int64_t tiles_n = (N + BN - 1) / BN;
    const int64_t bid = xcd_remap(blockIdx.x, tiles_m * tiles_n);
    const int64_t per_group = (int64_t)group_m * tiles_n;
    const int64_t grp = bid / per_group;
    const int64_t first_m = grp * group_m;
    const int64_t gsize = min((int64_t)group_m, tiles_m - first_m);
    const int64_t in_group = bid - grp * per_group;
    const int64_t m0 = (first_m + in_group % gsize) * BM;
    const int64_t n0 = (in_group / gsize) * BN;
    gemm_tile<BM, BN, TA, TB, VEC>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, vec_c, m0, n0, sm);
}

// ------------------------------------ GEMM with direct global-to-LDS loads
// M and N multiples of 128, K a multiple of 16, 16-byte aligned operands with
// even leading dimensions; any transposes.  128 x 128 tiles at two
// workgroups per CU (two waves per SIMD: 128 accumulator registers per
// lane), so that one workgroup's barrier, fragment reads and load waits
// overlap the other's MFMAs; the operand tiles go from global memory
// straight into LDS (global_load_lds_dwordx4: lane l of a wave instruction
// writes the 16 bytes at the instruction's LDS base + 16 l), with no
// register staging.  LDS layouts per stage and operand:
//   k-contiguous operand (A 'T', B 'N'): row r holds its eight k-pairs
//     (128 bytes), pair kp in slot kp ^ (r & 7); the global addresses of the
//     loads apply that permutation, so eight lanes still read one 128-byte
//     row.  A fragment is one 16-byte read.
//   row-contiguous operand (A 'N', B 'T'): k-row k holds the tile's 128 rows
//     (one 1 KB wave load), k-rows RC_LD doubles apart (the 64-byte pad puts
//     the four k-rows that one fragment read touches on alternate halves of
//     the banks).  A fragment is two 8-byte reads (k = 2 kp and 2 kp + 1).
// The fragments (lane g supplies the pair kp = 4 ks + g of its permuted
// row), the MFMA order and the epilogue are gemm_tile's, so the results are
// bit-identical to dgemm_mfma_kernel's.
constexpr int DL_TILE = 128;
constexpr int KC_SIZE = DL_TILE * BK;          // doubles of a k-contiguous operand tile
constexpr int RC_LD = DL_TILE + 8;             // doubles between the k-rows of a row-contiguous tile
constexpr int RC_SIZE = BK * RC_LD;
// s_waitcnt immediate (gfx9 fields: vmcnt bits 3:0 and 15:14, expcnt 6:4,
// lgkmcnt 11:8) waiting for vmcnt = 0 only; OR a count into bits 3:0.
constexpr int WAIT_VMCNT0 = 0x0F70;

__device__ __forceinline__ void lds_dma16(const double *g, double *lds) {
    __builtin_amdgcn_global_load_lds(static_cast<const void *>(g),
                                     (__attribute__((address_space(3))) void *)lds, 16, 0, 0);
}

// One operand tile in LDS: the loads of step k0 for wave `wave`, and the
// fragment of (row r, k-pair kp).
template <bool KCONTIG>
struct DlOperand {
    static constexpr int SIZE = KCONTIG ? KC_SIZE : RC_SIZE;
    const double *p;   // this thread's load address for k0 = 0
    int64_t ld;
    __device__ __forceinline__ DlOperand(const double *x, int64_t ld_, int64_t r0) : ld(ld_) {
        const int t = threadIdx.x, lane = t & 63;
        if (KCONTIG) {
            // slot q * 256 + t is row 32 q + t / 8, slot t % 8, holding the
            // pair (t % 8) ^ (row & 7); row & 7 does not depend on q
            const int lrow = t >> 3, lpair = (t & 7) ^ (lrow & 7);
            p = x + (r0 + lrow) * ld + 2 * lpair;
        } else {
            p = x + r0 + 2 * lane;
        }
    }
    __device__ __forceinline__ void issue(int64_t k0, double *tile) const {
        const int wave = threadIdx.x >> 6;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            if (KCONTIG) lds_dma16(p + q * 32 * ld + k0, tile + 2 * (q * 256 + 64 * wave));
            else lds_dma16(p + (k0 + 4 * q + wave) * ld, tile + (4 * q + wave) * RC_LD);
        }
    }
    __device__ __forceinline__ static f64x2 frag(const double *tile, int r, int kp) {
        if (KCONTIG) return *reinterpret_cast<const f64x2 *>(tile + 2 * (r * KP + (kp ^ (r & 7))));
        return f64x2{tile[2 * kp * RC_LD + r], tile[(2 * kp + 1) * RC_LD + r]};
    }
};

template <bool AK, bool BKC>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(2, 2))) void dgemm_dl_kernel(
    int64_t M, int64_t N, int64_t K, double alpha,
    const double *__restrict__ A, int64_t lda,
    const double *__restrict__ B, int64_t ldb,
    double beta, double *__restrict__ C, int64_t ldc,
    int vec_c, int group_m) {
    using OA = DlOperand<AK>;
    using OB = DlOperand<BKC>;
    constexpr int STAGE = OA::SIZE + OB::SIZE;
    __shared__ __attribute__((aligned(16))) double sm[2][STAGE];   // 64-68 KB
    constexpr int FM = 4, FN = 4;
    const int64_t tiles_m = M / DL_TILE, tiles_n = N / DL_TILE;
    const int64_t bid = xcd_remap(blockIdx.x, tiles_m * tiles_n);
    const int64_t per_group = (int64_t)group_m * tiles_n;
    const int64_t grp = bid / per_group;
    const int64_t first_m = grp * group_m;
    const int64_t gsize = min((int64_t)group_m, tiles_m - first_m);
    const int64_t in_group = bid - grp * per_group;
    const int64_t m0 = (first_m + in_group % gsize) * DL_TILE;
    const int64_t n0 = (in_group / gsize) * DL_TILE;

    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    const int g = lane >> 4, li = lane & 15;
    const int abase = wm * 64 + 4 * (li & 3) + (li >> 2);
    const int bbase = wn * 64 + li;
    const OA oa(A, lda, m0);
    const OB ob(B, ldb, n0);
    const int64_t nk = K / BK;
    auto issue = [&](int64_t kt, int st) {
        oa.issue(kt * BK, sm[st]);
        ob.issue(kt * BK, sm[st] + OA::SIZE);
    };

    f64x4 acc[FM][FN];
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j)
            acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(0.0, 0.0, f64x4{0.0, 0.0, 0.0, 0.0}, 0, 0, 0);
    f64x2 af[2][FM], bf[2][FN];
    auto read_frags = [&](int st, int ks, int h) {
        const int kp = 4 * ks + g;
#pragma unroll
        for (int i = 0; i < FM; ++i) af[h][i] = OA::frag(sm[st], abase + 16 * i, kp);
#pragma unroll
        for (int j = 0; j < FN; ++j) bf[h][j] = OB::frag(sm[st] + OA::SIZE, bbase + 16 * j, kp);
    };
    // MFMAs [from, to) of sub-step h, in gemm_tile's order: the .x block
    // (k = 2 kp), then the .y block (k = 2 kp + 1), each row-major in (i, j)
    auto mfma_range = [&](int h, int from, int to) {
#pragma unroll
        for (int q = 0; q < 2 * FM * FN; ++q) {
            if (q < from || q >= to) continue;
            const int i = (q % (FM * FN)) / FN, j = q % FN;
            if (q < FM * FN) acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[h][i].x, bf[h][j].x, acc[i][j], 0, 0, 0);
            else acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[h][i].y, bf[h][j].y, acc[i][j], 0, 0, 0);
        }
    };
    // Entering iteration kt: stage kt & 1 holds step kt and F[0] its first
    // sub-step; the loads of step kt + 1 into the other stage are in flight.
    // Iteration: read sub-step 1 into F[1]; MFMAs of F[0]; wait for this
    // wave's loads (step kt + 1) and barrier (every wave's loads have landed
    // and every wave has read step kt); refill stage kt & 1 with step kt + 2
    // (the last steps reload step nk - 1, never read: no branch); read the
    // first sub-step of step kt + 1 into F[0]; MFMAs of F[1].  The MFMAs of
    // F[1] after the barrier cover the latency of the F[0] reads, and those
    // of F[0] before it the latency of the F[1] reads, so the barrier's own
    // LDS wait finds nothing outstanding.  The scheduling barriers hold that
    // order: left alone the compiler moved the MFMAs past the barrier and
    // the fragment reads up to it, so that the barrier waited for the reads
    // with the MFMA pipe idle.  With both operands k-contiguous (A 'T',
    // B 'N') the F[1] reads measured fastest issued as one batch ahead of
    // all F[0] MFMAs (+1.5% against the compiler's order at 8192^3 and
    // 16384^3), for the other layouts behind the first four F[0] MFMAs
    // (+0.2-0.5%; the batch form there 0.3% slower).
    auto iteration = [&](int64_t kt, int st) {
        constexpr int HEAD = (AK && BKC) ? 0 : 4;
        mfma_range(0, 0, HEAD);
        __builtin_amdgcn_sched_barrier(0);
        read_frags(st, 1, 1);
        __builtin_amdgcn_sched_barrier(0);
        mfma_range(0, HEAD, 2 * FM * FN);
        __builtin_amdgcn_sched_barrier(0);
        __builtin_amdgcn_s_waitcnt(WAIT_VMCNT0);
        __syncthreads();
        issue(min(kt + 2, nk - 1), st);
        read_frags(st ^ 1, 0, 0);
        __builtin_amdgcn_sched_barrier(0);
        mfma_range(1, 0, 2 * FM * FN);
        __builtin_amdgcn_sched_barrier(0);
    };
    issue(0, 0);
    issue(min((int64_t)1, nk - 1), 1);
    __builtin_amdgcn_s_waitcnt(WAIT_VMCNT0 | 8);   // vmcnt(8): step 0 landed
    __syncthreads();
    read_frags(0, 0, 0);
    int64_t kt = 0;
    for (; kt + 1 < nk; kt += 2) {
        iteration(kt, 0);
        iteration(kt + 1, 1);
    }
    if (kt < nk) iteration(kt, 0);
    __builtin_amdgcn_s_waitcnt(WAIT_VMCNT0);   // no LDS writes outstanding at exit

    // ---- epilogue (gemm_tile's): lane holds rows 4g..4g+3 of column li
    const bool beta0 = (beta == 0.0);
#pragma unroll
    for (int i = 0; i < FM; ++i) {
        const int64_t row = m0 + wm * 64 + 16 * i + 4 * g;
#pragma unroll
        for (int j = 0; j < FN; ++j) {
            const int64_t col = n0 + wn * 64 + 16 * j + li;
            double *cp = C + row + col * ldc;
            const f64x4 r = acc[i][j];
            if (vec_c) {
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
                    double val = alpha * r[e];
                    if (!beta0) val = fma(beta, cp[e], val);
                    cp[e] = val;
                }
            }
        }
    }
}

// Grouped tiles without LDS ("direct"): one wave per (16 FI) x (16 FJ) C tile
// (64 x 64 or 32 x 32: up to FI x FJ blocks of 16 x 16 in AGPR accumulators;
// blocks outside C are skipped),
// fragments loaded straight from global memory (L1 / L2 cached) one 16-deep
// k-step ahead of their MFMAs; no LDS and no barriers, so small problems pay
// neither the staging round trip nor the workgroup synchronisation.  The
// order of the products in every accumulator is gemm_tile's (per 16-deep
// step: k = 2g, then 2g + 1, then 8 + 2g, then 9 + 2g in lane group g, zero
// beyond K), the A rows are permuted the same way and the epilogue is the
// same, so the results are bit-identical to gemm_tile's.  The 32 x 32 shape
// pads small problems less (0.77 of the padded MFMA work is useful on the
// 4096-problem log-uniform batch, against 0.63 for 64 x 64) and, with a
// quarter of the accumulators, leaves room for several waves per SIMD to hide
// the load latency.
template <int FI, int FJ, bool TA, bool TB>
__device__ __forceinline__ void direct_tile(const GroupedProblem &q, int64_t m0, int64_t n0) {
    const int lane = threadIdx.x & 63, g = lane >> 4, li = lane & 15;
    const int arow = 4 * (li & 3) + (li >> 2);
    const int64_t M = q.m, N = q.n;
    const bool scale_only = (q.k == 0 || q.alpha == 0.0);
    const int64_t K = scale_only ? 0 : q.k;
    const double alpha = scale_only ? 0.0 : q.alpha;
    const int rows_in = (int)min((int64_t)FI, (M - m0 + 15) / 16);
    const int cols_in = (int)min((int64_t)FJ, (N - n0 + 15) / 16);
    const double *__restrict__ A = q.a;
    const double *__restrict__ B = q.b;
    f64x4 acc[FI][FJ];
#pragma unroll
    for (int i = 0; i < FI; ++i)
#pragma unroll
        for (int j = 0; j < FJ; ++j)
            acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(0.0, 0.0, f64x4{0.0, 0.0, 0.0, 0.0}, 0, 0, 0);
    const int koff[4] = {2 * g, 2 * g + 1, 8 + 2 * g, 9 + 2 * g};
    auto load_step = [&](int64_t k0, double (&a)[FI][4], double (&b)[FJ][4]) {
#pragma unroll
        for (int i = 0; i < FI; ++i) {
            const int64_t r = m0 + 16 * i + arow;
#pragma unroll
            for (int x = 0; x < 4; ++x) {
                const int64_t k = k0 + koff[x];
                a[i][x] = (i < rows_in && r < M && k < K) ? (TA ? A[k + r * q.lda] : A[r + k * q.lda]) : 0.0;
            }
        }
#pragma unroll
        for (int j = 0; j < FJ; ++j) {
            const int64_t c = n0 + 16 * j + li;
#pragma unroll
            for (int x = 0; x < 4; ++x) {
                const int64_t k = k0 + koff[x];
                b[j][x] = (j < cols_in && c < N && k < K) ? (TB ? B[c + k * q.ldb] : B[k + c * q.ldb]) : 0.0;
            }
        }
    };
    auto mfma_step = [&](const double (&a)[FI][4], const double (&b)[FJ][4]) {
#pragma unroll
        for (int x = 0; x < 4; ++x)
#pragma unroll
            for (int i = 0; i < FI; ++i)
#pragma unroll
                for (int j = 0; j < FJ; ++j)
                    if (i < rows_in && j < cols_in)
                        acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i][x], b[j][x], acc[i][j], 0, 0, 0);
    };
    const int64_t nsteps = (K + 15) / 16;
    double a0[FI][4], b0[FJ][4], a1[FI][4], b1[FJ][4];
    if (nsteps > 0) load_step(0, a0, b0);
    for (int64_t s = 0; s < nsteps; s += 2) {
        if (s + 1 < nsteps) load_step(16 * (s + 1), a1, b1);
        mfma_step(a0, b0);
        if (s + 2 < nsteps) load_step(16 * (s + 2), a0, b0);
        if (s + 1 < nsteps) mfma_step(a1, b1);
    }
    // ---- epilogue (gemm_tile's): lane holds rows 4g..4g+3 of column li
    const bool beta0 = (q.beta == 0.0);
    const double beta = q.beta;
#pragma unroll
    for (int i = 0; i < FI; ++i) {
        const int64_t row = m0 + 16 * i + 4 * g;
        if (row >= M) continue;
#pragma unroll
        for (int j = 0; j < FJ; ++j) {
            const int64_t col = n0 + 16 * j + li;
            if (col >= N) continue;
            double *cp = q.c + row + col * q.ldc;
            const f64x4 r = acc[i][j];
            if (q.vec_c && row + 3 < M) {
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

// Grouped launches walk virtual tile ids v = blockIdx.x + i * gridDim.x; the
// hardware places workgroup b on XCD b % 8.  Consecutive tiles (of one
// problem, which share A and B panels) are kept on one XCD and its L2 by
// handing out chunks of GROUP_CHUNK consecutive tiles: XCD x takes chunks x,
// x + 8, x + 16, ... in dispatch order, so the decreasing-cost order of the
// problems is kept at chunk granularity.  Bijective on [0, T): ids beyond
// the last whole round of 8 chunks map to themselves.
constexpr int64_t GROUP_CHUNK = 8;
__device__ __forceinline__ int64_t grouped_tile_of(int64_t v, int64_t T) {
    constexpr int64_t span = NXCD * GROUP_CHUNK;
    if (v >= T / span * span) return v;
    const int64_t xcd = v % NXCD, local = v / NXCD;
    return ((local / GROUP_CHUNK) * NXCD + xcd) * GROUP_CHUNK + local % GROUP_CHUNK;
}

// The descriptor array is followed by the problems' first tile indices
// (int64, same order): the binary search for a tile's problem walks this
// compact array (8 bytes per problem, so its upper levels stay in the scalar
// cache) instead of one 112-byte descriptor per step.
__device__ __forceinline__ int grouped_owner(const GroupedProblem *probs, int nprob, int64_t t) {
    const int64_t *starts = reinterpret_cast<const int64_t *>(probs + nprob);
    int lo = 0, hi = nprob - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (starts[mid] <= t) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

template <int FI, int FJ>
__global__ __launch_bounds__(64) void dgemm_grouped_direct(const GroupedProblem *__restrict__ probs, int nprob,
                                                           int64_t total_tiles) {
    for (int64_t v = blockIdx.x; v < total_tiles; v += gridDim.x) {
        const int64_t t = grouped_tile_of(v, total_tiles);
        const GroupedProblem &q = probs[grouped_owner(probs, nprob, t)];
        const int64_t local = t - q.tile_start;
        const int64_t tiles_m = (q.m + 16 * FI - 1) / (16 * FI);
        const int64_t m0 = (local % tiles_m) * (16 * FI), n0 = (local / tiles_m) * (16 * FJ);
        switch (q.flags & 3) {   // bit 0: A transposed, bit 1: B transposed
        case 0: direct_tile<FI, FJ, false, false>(q, m0, n0); break;
        case 1: direct_tile<FI, FJ, true, false>(q, m0, n0); break;
        case 2: direct_tile<FI, FJ, false, true>(q, m0, n0); break;
        default: direct_tile<FI, FJ, true, true>(q, m0, n0); break;
        }
    }
}

// C <- beta * C; beta == 0 writes zeros without reading C.
// Elementwise kernels over an M x N column-major block: threads run down the
// rows of a column (coalesced), blocks stride over the columns.
__global__ void scale_kernel(int64_t M, int64_t N, double beta, double *__restrict__ C, int64_t ldc) {
    const int64_t row = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (row >= M) return;
    for (int64_t col = blockIdx.y; col < N; col += gridDim.y) {
        double *p = C + row + col * ldc;
        *p = (beta == 0.0) ? 0.0 : beta * *p;
    }
}

// C += beta * X over an M x N block (X leading dimension ldx).
__global__ void add_scaled_kernel(int64_t M, int64_t N, double beta, const double *__restrict__ X,
                                  int64_t ldx, double *__restrict__ C, int64_t ldc) {
    const int64_t row = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (row >= M) return;
    for (int64_t col = blockIdx.y; col < N; col += gridDim.y) C[row + col * ldc] += beta * X[row + col * ldx];
}

inline dim3 elementwise_grid(int64_t m, int64_t n) {
    return dim3((unsigned)((m + 255) / 256), (unsigned)std::min<int64_t>(n, 2048));
}

inline bool aligned16(const void *p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; }

template <int BM, int BN>
int launch_cfg(char ta, char tb, int64_t m, int64_t n, int64_t k, double alpha,
               const double *a, int64_t lda, const double *b, int64_t ldb,
               double beta, double *c, int64_t ldc, int group_m, hipStream_t s) {
    const int64_t tiles = ((m + BM - 1) / BM) * ((n + BN - 1) / BN);
    if (tiles > 0x7fffffffLL) {
        set_error("dgemm: %lld output tiles exceed the grid limit", (long long)tiles);
        return COSMA_AMD_ERR_UNSUPPORTED;
    }
    const bool TA = ta == 'T', TB = tb == 'T';
    // 16-byte operand loads need aligned pointers, even leading dimensions
    // and, for row-contiguous operands, an even row count.
    const bool vec = aligned16(a) && (lda % 2 == 0) && aligned16(b) && (ldb % 2 == 0) &&
                     (TA || m % 2 == 0) && (!TB || n % 2 == 0);
    const int vec_c = aligned16(c) && (ldc % 2 == 0);
    const dim3 grid((unsigned)tiles), block(NT);
#define CA_LAUNCH(X, Y, V)                                                                        \
    hipLaunchKernelGGL((dgemm_mfma_kernel<BM, BN, X, Y, V>), grid, block, 0, s, m, n, k, alpha, a, \
                       lda, b, ldb, beta, c, ldc, vec_c, group_m)
    if (vec) {
        if (!TA && !TB) CA_LAUNCH(false, false, true);
        else if (TA && !TB) CA_LAUNCH(true, false, true);
        else if (!TA && TB) CA_LAUNCH(false, true, true);
        else CA_LAUNCH(true, true, true);
    } else {
        if (!TA && !TB) CA_LAUNCH(false, false, false);
        else if (TA && !TB) CA_LAUNCH(true, false, false);
        else if (!TA && TB) CA_LAUNCH(false, true, false);
        else CA_LAUNCH(true, true, false);
    }
#undef CA_LAUNCH
    CA_HIP(hipGetLastError());
    return COSMA_AMD_OK;
}

// ------------------------------------------------------------ grouped GEMM
// One persistent launch over the tiles of many independent problems.  Tiles
// are numbered problem by problem (problems in the order of the descriptor
// array, which the host sorts by decreasing cost) and column-major inside a
// problem; workgroup w takes the virtual ids w, w + G, w + 2G, ... mapped to
// tiles by grouped_tile_of (XCD-local chunks).  Each tile runs the
// same gemm_tile as the single-GEMM kernel, so a problem's result is
// bit-identical to a cosma_amd_dgemm call on it.
template <int BM, int BN>
__global__ __launch_bounds__(NT) void dgemm_grouped_kernel(const GroupedProblem *__restrict__ probs, int nprob,
                                                           int64_t total_tiles) {
    __shared__ TileSmem<BM, BN> sm;
    for (int64_t v = blockIdx.x; v < total_tiles; v += gridDim.x) {
        const int64_t t = grouped_tile_of(v, total_tiles);
        // problem owning tile t: last p with tile_start <= t
        const int lo = grouped_owner(probs, nprob, t);
        const GroupedProblem &q = probs[lo];
        const int64_t local = t - q.tile_start;
        const int64_t tiles_m = (q.m + BM - 1) / BM;
        const int64_t m0 = (local % tiles_m) * BM, n0 = (local / tiles_m) * BN;
        const bool scale_only = (q.k == 0 || q.alpha == 0.0);
        const int64_t K = scale_only ? 0 : q.k;
        const double alpha = scale_only ? 0.0 : q.alpha;
        __syncthreads();   // the previous tile's last LDS reads are done
#define CA_TILE(X, Y, V)                                                                              \
    gemm_tile<BM, BN, X, Y, V, true>(q.m, q.n, K, alpha, q.a, q.lda, q.b, q.ldb, q.beta, q.c, q.ldc, q.vec_c, m0, n0, sm)
        const int f = q.flags;   // bit 0: A transposed, bit 1: B transposed, bit 2: 16-byte loads
        switch (f) {
        case 0: CA_TILE(false, false, false); break;
        case 1: CA_TILE(true, false, false); break;
        case 2: CA_TILE(false, true, false); break;
        case 3: CA_TILE(true, true, false); break;
        case 4: CA_TILE(false, false, true); break;
        case 5: CA_TILE(true, false, true); break;
        case 6: CA_TILE(false, true, true); break;
        default: CA_TILE(true, true, true); break;
        }
#undef CA_TILE
    }
}

}  // namespace

int launch_scale(int64_t m, int64_t n, double beta, double *c, int64_t ldc, hipStream_t s) {
    if (m == 0 || n == 0 || beta == 1.0) return COSMA_AMD_OK;
    hipLaunchKernelGGL(scale_kernel, elementwise_grid(m, n), dim3(256), 0, s, m, n, beta, c, ldc);
    CA_HIP(hipGetLastError());
    return COSMA_AMD_OK;
}

int launch_add_scaled(int64_t m, int64_t n, double beta, const double *x, int64_t ldx, double *c,
                      int64_t ldc, hipStream_t s) {
    if (m == 0 || n == 0 || beta == 0.0) return COSMA_AMD_OK;
    hipLaunchKernelGGL(add_scaled_kernel, elementwise_grid(m, n), dim3(256), 0, s, m, n, beta, x, ldx, c, ldc);
    CA_HIP(hipGetLastError());
    return COSMA_AMD_OK;
}

// The direct-LDS kernel's conditions (see dgemm_dl_kernel).
bool dl_eligible(int64_t m, int64_t n, int64_t k, const double *a, int64_t lda, const double *b, int64_t ldb) {
    return m % DL_TILE == 0 && n % DL_TILE == 0 && k % BK == 0 && k > 0 && aligned16(a) && aligned16(b) &&
           lda % 2 == 0 && ldb % 2 == 0 && (m / DL_TILE) * (n / DL_TILE) <= 0x7fffffffLL;
}

int launch_dl(char ta, char tb, int64_t m, int64_t n, int64_t k, double alpha, const double *a, int64_t lda,
              const double *b, int64_t ldb, double beta, double *c, int64_t ldc, int group_m, hipStream_t s) {
    const int vec_c = aligned16(c) && (ldc % 2 == 0);
    const dim3 grid((unsigned)((m / DL_TILE) * (n / DL_TILE))), block(NT);
#define CA_LAUNCH_DL(X, Y)                                                                                 \
    hipLaunchKernelGGL((dgemm_dl_kernel<X, Y>), grid, block, 0, s, m, n, k, alpha, a, lda, b, ldb, beta, c, \
                       ldc, vec_c, group_m)
    // template flags: operand k-contiguous (A 'T', B 'N')
    if (ta == 'T' && tb == 'N') CA_LAUNCH_DL(true, true);
    else if (ta == 'N' && tb == 'N') CA_LAUNCH_DL(false, true);
    else if (ta == 'N' && tb == 'T') CA_LAUNCH_DL(false, false);
    else CA_LAUNCH_DL(true, false);
#undef CA_LAUNCH_DL
    CA_HIP(hipGetLastError());
    return COSMA_AMD_OK;
}

// Tile configuration: 0 = automatic, 1 = 128x128, 2 = 256x128, 3 = 128x256,
// 4 = the direct-LDS kernel (128x128, two workgroups per CU) where its
// conditions hold, else as 0.
static int g_cfg_override = 0;
static int g_group_override = 0;

int launch_dgemm_cfg(int cfg, int group_m, char ta, char tb, int64_t m, int64_t n, int64_t k,
                     double alpha, const double *a, int64_t lda, const double *b, int64_t ldb,
                     double beta, double *c, int64_t ldc, hipStream_t s) {
    if (m == 0 || n == 0) return COSMA_AMD_OK;
    if (k == 0 || alpha == 0.0) return launch_scale(m, n, beta, c, ldc, s);
    // Large tiles need enough of them to fill 256 CUs; of the two large
    // shapes 128 x 256 measured faster for NN and TN at 8192^3 and 16384^3.
    const int64_t t_large = ((m + 127) / 128) * ((n + 255) / 256);
    const int auto_cfg = (t_large >= 512) ? 3 : 1;
    if (cfg == 4 || cfg == 0) {
        // The direct-LDS kernel (bit-identical results) against 128 x 256
        // tiles at 8192^3 / 16384^3 measured +5% / +4% (NN), +3% / +2.7%
        // (TN), +6% / +6% (NT), +3% / +3% (TT).  With 512 tiles resident (two
        // per CU) against 256, its last, partial round can waste more: it is
        // taken automatically from 1024 tiles on, unless the other choice's
        // last round is fuller by more than its 4% advantage.
        bool take = cfg == 4;
        if (!take && dl_eligible(m, n, k, a, lda, b, ldb)) {
            const int64_t t = (m / DL_TILE) * (n / DL_TILE);
            const int64_t t_alt = auto_cfg == 3 ? t_large : ((m + 127) / 128) * ((n + 127) / 128);
            auto fill = [](int64_t tiles, int64_t slots) {
                return (double)tiles / (double)(((tiles + slots - 1) / slots) * slots);
            };
            take = t >= 1024 && 1.04 * fill(t, 512) >= fill(t_alt, 256);
        }
        if (take && dl_eligible(m, n, k, a, lda, b, ldb))
            return launch_dl(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, group_m > 0 ? group_m : 8, s);
        cfg = 0;
    }
    if (cfg == 0) cfg = auto_cfg;
    switch (cfg) {
    case 1: return launch_cfg<128, 128>(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, group_m > 0 ? group_m : 8, s);
    case 2: return launch_cfg<256, 128>(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, group_m > 0 ? group_m : 4, s);
    case 3: return launch_cfg<128, 256>(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, group_m > 0 ? group_m : 8, s);
    default:
        set_error("dgemm: unknown tile configuration %d", cfg);
        return COSMA_AMD_ERR_INVALID_ARG;
    }
}

int launch_dgemm(char ta, char tb, int64_t m, int64_t n, int64_t k, double alpha,
                 const double *a, int64_t lda, const double *b, int64_t ldb,
                 double beta, double *c, int64_t ldc, hipStream_t s) {
    return launch_dgemm_cfg(g_cfg_override, g_group_override, ta, tb, m, n, k, alpha, a, lda, b,
                            ldb, beta, c, ldc, s);
}

void set_dgemm_tuning(int cfg, int group_m) {
    g_cfg_override = cfg;
    g_group_override = group_m;
}

int grouped_tile_shape(int cfg, int *bm, int *bn) {
    switch (cfg) {
    case 1: *bm = 64; *bn = 64; return COSMA_AMD_OK;
    case 2: *bm = 128; *bn = 64; return COSMA_AMD_OK;
    case 3: *bm = 128; *bn = 128; return COSMA_AMD_OK;
    case 4: *bm = 64; *bn = 64; return COSMA_AMD_OK;   // direct: one wave per tile, no LDS
    case 5: *bm = 32; *bn = 32; return COSMA_AMD_OK;   // direct, 32 x 32
    default:
        set_error("grouped dgemm: unknown tile configuration %d", cfg);
        return COSMA_AMD_ERR_INVALID_ARG;
    }
}

int grouped_problem_flags(char ta, char tb, int64_t m, int64_t n, const double *a, int64_t lda, const double *b,
                          int64_t ldb, const double *c, int64_t ldc, int *flags, int *vec_c) {
    const bool TA = ta == 'T', TB = tb == 'T';
    const bool vec = aligned16(a) && (lda % 2 == 0) && aligned16(b) && (ldb % 2 == 0) && (TA || m % 2 == 0) &&
                     (!TB || n % 2 == 0);
    *flags = (TA ? 1 : 0) | (TB ? 2 : 0) | (vec ? 4 : 0);
    *vec_c = aligned16(c) && (ldc % 2 == 0);
    return COSMA_AMD_OK;
}

int launch_grouped(int cfg, const GroupedProblem *probs, int nprob, int64_t total_tiles, int grid, hipStream_t s) {
    if (nprob == 0 || total_tiles == 0) return COSMA_AMD_OK;
    const dim3 g((unsigned)grid), b(NT);
    switch (cfg) {
    case 1: hipLaunchKernelGGL((dgemm_grouped_kernel<64, 64>), g, b, 0, s, probs, nprob, total_tiles); break;
    case 2: hipLaunchKernelGGL((dgemm_grouped_kernel<128, 64>), g, b, 0, s, probs, nprob, total_tiles); break;
    case 3: hipLaunchKernelGGL((dgemm_grouped_kernel<128, 128>), g, b, 0, s, probs, nprob, total_tiles); break;
    case 4: hipLaunchKernelGGL((dgemm_grouped_direct<4, 4>), g, dim3(64), 0, s, probs, nprob, total_tiles); break;
    case 5: hipLaunchKernelGGL((dgemm_grouped_direct<2, 2>), g, dim3(64), 0, s, probs, nprob, total_tiles); break;
    default:
        set_error("grouped dgemm: unknown tile configuration %d", cfg);
        return COSMA_AMD_ERR_INVALID_ARG;
    }
    CA_HIP(hipGetLastError());
    return COSMA_AMD_OK;
}

}  // namespace cosma_amd
