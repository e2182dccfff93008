// Batched INT8 MFMA GEMM with pair lists and residue / FP64 epilogues.
// See i8gemm.hpp for the contract and the tile layout.
#include <algorithm>
#include <type_traits>

#include "i8gemm.hpp"

namespace cosma_amd {
namespace {

constexpr int BM = 256, BN = 256, BK = 128, NT = 256;
constexpr int NXCD = 8;
constexpr int GROUP_M = 4;

__device__ __forceinline__ int64_t xcd_remap(int64_t b, int64_t T) {
    const int64_t xcd = b % NXCD, local = b / NXCD;
    const int64_t q = T / NXCD, r = T % NXCD;
    return (xcd < r) ? xcd * (q + 1) + local : r * (q + 1) + (xcd - r) * q + local;
}

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

// One MFMA on 16-byte fragments: INT8 (64 k, INT32 sums) or BF16 (32 k,
// FP32 sums).
template <bool BF16>
__device__ __forceinline__ auto mma(const i32x4 &a, const i32x4 &b, const std::conditional_t<BF16, f32x4, i32x4> &c) {
    if constexpr (BF16)
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c,
                                                       0, 0, 0);
    else
        return __builtin_amdgcn_mfma_i32_16x16x64_i8(a, b, c, 0, 0, 0);
}

// WIDE (EPI_RES8 with INT8 operands and K > 32768): the residue reduction
// for sums of any magnitude (a template parameter so that each kernel holds
// one form: both forms in one kernel made the register allocator spill).
template <int EPI, bool BF16, bool WIDE = false>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(1, 1))) void i8gemm_kernel(I8Problem p, I8Jobs jobs) {
    using acc_t = std::conditional_t<BF16, f32x4, i32x4>;
    // LDS: two stages of 256 x 128-byte rows per operand (128 KB in all).
    // Row r holds its eight 16-byte chunks in the order chunk ^ (r & 7): the
    // stores (eight consecutive lanes per row, one full 128-byte cache line
    // per row from global memory) fill a contiguous 1 KB per wave
    // instruction, and each group of eight lanes of a fragment read (rows
    // r..r+7, one chunk) touches eight different 16-byte bank groups.
    __shared__ __attribute__((aligned(16))) i32x4 sa[2][BM * BK / 16];
    __shared__ __attribute__((aligned(16))) i32x4 sb[2][BN * BK / 16];

    const int z = blockIdx.y;
    const int64_t tiles_m = (p.M + BM - 1) / BM, tiles_n = (p.N + BN - 1) / BN;
    const int64_t bid = xcd_remap(blockIdx.x, tiles_m * tiles_n);
    const int64_t per_group = (int64_t)GROUP_M * tiles_n;
    const int64_t grp = bid / per_group;
    const int64_t first_m = grp * GROUP_M;
    const int64_t gsize = min((int64_t)GROUP_M, tiles_m - first_m);
    const int64_t in_group = bid - grp * per_group;
    const int64_t m0 = (first_m + in_group % gsize) * BM;
    const int64_t n0 = (in_group / gsize) * BN;

    const int t = threadIdx.x;
    const int lane = t & 63, wave = t >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    const int g = lane >> 4, li = lane & 15;

    // accumulators zeroed by an MFMA so they live in AGPRs across the loop
    acc_t acc[8][8];
    const i32x4 zero4 = {0, 0, 0, 0};
    const acc_t zacc = {0, 0, 0, 0};
#pragma unroll
    for (int i = 0; i < 8; ++i)
#pragma unroll
        for (int j = 0; j < 8; ++j) acc[i][j] = mma<BF16>(zero4, zero4, zacc);

    const int ksteps = (int)(p.K / BK);
    const int nsteps = jobs.npairs[z] * ksteps;
    // slice indices of the pairs, held in registers (byte p of the word = pair p)
    uint64_t pa_bits = 0, pb_bits = 0;
#pragma unroll
    for (int q = 0; q < I8_MAX_PAIRS; ++q) {
        pa_bits |= (uint64_t)jobs.pa[z][q] << (8 * q);
        pb_bits |= (uint64_t)jobs.pb[z][q] << (8 * q);
    }

    // global -> register staging: 8 x 16 bytes of A and of B per thread and
    // step.  Addresses are a wave-uniform base (slice, tile, k0) plus a 32-bit
    // per-thread offset (row (t >> 3) + 32 q, chunk t & 7), so each load needs
    // one offset register instead of a 64-bit address.
    i32x4 ra[8], rb[8];
    const uint32_t lda32 = (uint32_t)p.lda, ldb32 = (uint32_t)p.ldb;
    const uint32_t offa = (uint32_t)(t >> 3) * lda32 + (uint32_t)(t & 7) * 16u;
    const uint32_t offb = (uint32_t)(t >> 3) * ldb32 + (uint32_t)(t & 7) * 16u;
    auto load_step = [&](int step) {
        const int pair = step / ksteps;
        const int64_t k0 = (int64_t)(step - pair * ksteps) * BK;
        const int8_t *A = p.a + (int64_t)((pa_bits >> (8 * pair)) & 0xff) * p.a_stride + m0 * p.lda + k0;
        const int8_t *B = p.b + (int64_t)((pb_bits >> (8 * pair)) & 0xff) * p.b_stride + n0 * p.ldb + k0;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            ra[q] = *reinterpret_cast<const i32x4 *>(A + (uint64_t)(offa + (uint32_t)q * 32u * lda32));
            rb[q] = *reinterpret_cast<const i32x4 *>(B + (uint64_t)(offb + (uint32_t)q * 32u * ldb32));
        }
    };
    auto swz = [](int row, int chunk) { return row * 8 + (chunk ^ (row & 7)); };
    auto store_step = [&](int st) {
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int idx = t + NT * q;
            const int row = idx >> 3, chunk = idx & 7;
            sa[st][swz(row, chunk)] = ra[q];
            sb[st][swz(row, chunk)] = rb[q];
        }
    };

    // Fragments of the two 64-deep halves of a step: F[0] (bytes 0..63 of
    // the rows) and F[1] (bytes 64..127).
    i32x4 fa[2][8], fb[2][8];
    auto read_frags = [&](int st, int h) {
#pragma unroll
        for (int j = 0; j < 8; ++j) fb[h][j] = sb[st][swz(wn * 128 + 16 * j + li, 4 * h + g)];
#pragma unroll
        for (int i = 0; i < 8; ++i) fa[h][i] = sa[st][swz(wm * 128 + 16 * i + li, 4 * h + g)];
    };
    auto mfmas = [&](int h) {
#pragma unroll
        for (int i = 0; i < 8; ++i)
#pragma unroll
            for (int j = 0; j < 8; ++j)
                acc[i][j] = mma<BF16>(fa[h][i], fb[h][j], acc[i][j]);
    };
    // BF16: the exact FP32 sums are reduced modulo the job's modulus (see
    // i8gemm.hpp), x <- x - m * rint(x / m), exact on integers below 2^24.
    // Written as plain FP32 code the reduction made the compiler move the
    // accumulators out of the AGPRs (500+ spilled registers), so each element
    // is read, reduced and written back in place by a short asm block.  The
    // hazard recognizer does not look into asm, so the required wait states
    // are inserted by hand: s_nops before the first read (MFMA result ->
    // VALU read) and after the last write (VALU AGPR write -> MFMA SrcC read).
    const float finv_m = jobs.fmod[z][0], neg_m = jobs.fmod[z][1];
    auto reduce = [&]() {
        if constexpr (BF16) {
            asm volatile("s_nop 7\n\ts_nop 7\n\ts_nop 7" ::: "memory");
#pragma unroll
            for (int i = 0; i < 8; ++i)
#pragma unroll
                for (int j = 0; j < 8; ++j)
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        float x = acc[i][j][e], t, q;
                        asm volatile("v_accvgpr_read_b32 %1, %0\n\t"
                                     "v_mul_f32 %2, %1, %4\n\t"
                                     "v_rndne_f32 %2, %2\n\t"
                                     "v_fmac_f32 %1, %2, %3\n\t"
                                     "v_accvgpr_write_b32 %0, %1"
                                     : "+a"(x), "=&v"(t), "=&v"(q)
                                     : "v"(neg_m), "v"(finv_m));
                        acc[i][j][e] = x;
                    }
            asm volatile("s_nop 7\n\ts_nop 7" ::: "memory");
        }
    };
    const int last = nsteps - 1;
    // Step s lives in LDS stage s & 1.  Entering iteration s, the registers
    // hold the global data of step s + 1 and F[0] holds the first half of
    // step s.  Iteration s: read the second half of step s into F[1], run
    // both halves' MFMAs with the 16 LDS stores of step s + 1 (into the other
    // stage, whose last reads happened before the previous barrier) spread
    // one per 4 MFMAs over the first half and the 16 global loads of step
    // s + 2 one per 4 MFMAs over the second, barrier, read the first half of
    // step s + 1.  Left to itself the scheduler clusters all 16 stores (each
    // behind its load wait) right after the first MFMA, which stalls the
    // MFMA stream; the group barriers fix the interleave (11% faster at
    // 8192^3).  The body is branch-free (loads past the end re-read the last
    // step; the extra store and fragment reads are never consumed), so the
    // compiler's wait counts stay exact; nsteps is even (K is a multiple of
    // 256), so the stage indices are constants.  The plain scheduling
    // barriers keep each batch of fragment reads together ahead of the
    // MFMAs: left alone, the scheduler sinks every A-fragment read to just
    // before its first use and waits for it there.
    auto iteration = [&](int s, int st) {
        read_frags(st, 1);
        __builtin_amdgcn_sched_barrier(0);
        mfmas(0);
        mfmas(1);
        store_step(st ^ 1);
        load_step(min(s + 2, last));
#pragma unroll
        for (int w = 0; w < 16; ++w) {
            __builtin_amdgcn_sched_group_barrier(0x008, 4, 0);   // MFMA
            __builtin_amdgcn_sched_group_barrier(0x200, 1, 0);   // DS write
        }
#pragma unroll
        for (int w = 0; w < 16; ++w) {
            __builtin_amdgcn_sched_group_barrier(0x008, 4, 0);   // MFMA
            __builtin_amdgcn_sched_group_barrier(0x020, 1, 0);   // VMEM read
        }
        __builtin_amdgcn_sched_barrier(0);
        __syncthreads();
        read_frags(st ^ 1, 0);
        __builtin_amdgcn_sched_barrier(0);
    };

    if (nsteps > 0) {
        load_step(0);
        store_step(0);
        load_step(min(1, last));
        __syncthreads();
        read_frags(0, 0);
        __builtin_amdgcn_sched_barrier(0);
        for (int s = 0; s < nsteps; s += 2) {
            iteration(s, 0);
            iteration(s + 1, 1);
            if (BF16 && (s + 2) % BF16_FLUSH_STEPS == 0) reduce();
        }
    }

    // ---- epilogue: lane holds rows 4g..4g+3 (contiguous) of column li
    const double scale = jobs.scale[z];
    if constexpr (EPI == EPI_RES8) {
        // The 256 x 256 int8 residues go through LDS (stage 0 of A, 64 KB,
        // free once every wave has left the main loop) so that the global
        // writes are whole 256-byte column runs instead of 16-byte pieces in
        // 16 columns per instruction.  Column c of the tile occupies bytes
        // [256 c, 256 c + 256) with row r at r ^ (16 (c & 15)): the 4-byte
        // writes of a wave (16 columns x 4 row quads) hit 64 different banks,
        // and each 16-byte chunk stays a contiguous chunk.
        // sym_mod's quarter-rate multiply and range corrections were most
        // of the epilogue's instructions.  sym_mod_odd_bits needs sums below
        // 2^29 in magnitude, i.e. npairs * K <= 32768 (|x| <= npairs * K *
        // 2^14; the BF16 sums stay below 2^24 for every K), else
        // sym_mod_odd_bits_wide applies.
        // Both return a word whose low byte is x's residue also for m = 256
        // (c16 = 0: y = x & 0xffff, and y - q m = x mod 256 whatever q);
        // even moduli other than 256 are rejected at launch.
        uint8_t *tile = reinterpret_cast<uint8_t *>(&sa[0][0]);
        __syncthreads();
        auto stage = [&](auto residue_bits) {
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                const int rt = wm * 128 + 16 * i + 4 * g;
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    const int ct = wn * 128 + 16 * j + li;
                    uint32_t b[4];
#pragma unroll
                    for (int e = 0; e < 4; ++e) b[e] = residue_bits((int)acc[i][j][e]);   // BF16: exact integers
                    // low bytes of b[0..3] -> one word (v_perm_b32 picks bytes 0-3 of its
                    // second operand and 4-7 of its first; selector 0x0c gives zero)
                    const uint32_t lo = __builtin_amdgcn_perm(b[1], b[0], 0x0c0c0400u);
                    const uint32_t hi = __builtin_amdgcn_perm(b[3], b[2], 0x0c0c0400u);
                    *reinterpret_cast<uint32_t *>(tile + ct * 256 + (rt ^ (16 * (ct & 15)))) =
                        __builtin_amdgcn_perm(hi, lo, 0x05040100u);
                }
            }
        };
        const float c16 = jobs.fmod[z][2];
        if constexpr (WIDE)
            stage([&](int x) { return sym_mod_odd_bits_wide(x, c16, finv_m, neg_m); });
        else
            stage([&](int x) { return sym_mod_odd_bits(x, c16, finv_m, neg_m); });
        __syncthreads();
        int8_t *out = static_cast<int8_t *>(p.out) + (int64_t)z * p.out_stride;
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int id = t + NT * q, ct = id >> 4, c = id & 15;
            const int64_t col = n0 + ct, row = m0 + 16 * c;
            if (col >= p.N || row >= p.M) continue;
            const i32x4 v = *reinterpret_cast<const i32x4 *>(tile + ct * 256 + 16 * (c ^ (ct & 15)));
            int8_t *o = out + row + col * p.ldo;
            if (row + 16 <= p.M) {
                *reinterpret_cast<i32x4 *>(o) = v;
            } else {
                union { i32x4 q; int8_t b[16]; } u;
                u.q = v;
                for (int e = 0; e < 16 && row + e < p.M; ++e) o[e] = u.b[e];
            }
        }
    } else {
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int64_t row = m0 + wm * 128 + 16 * i + 4 * g;
            if (row >= p.M) continue;
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int64_t col = n0 + wn * 128 + 16 * j + li;
                if (col >= p.N) continue;
                const acc_t r = acc[i][j];
                double *o = static_cast<double *>(p.out) + (int64_t)z * p.out_stride + row + col * p.ldo;
                for (int e = 0; e < 4 && row + e < p.M; ++e) {
                    const double val = scale * (double)r[e];
                    o[e] = p.accumulate ? o[e] + val : val;
                }
            }
        }
    }
}

}  // namespace

int launch_i8gemm(const I8Problem &p, const I8Jobs &jobs, int Z, I8Epilogue epi, hipStream_t s) {
    if (p.M == 0 || p.N == 0 || Z == 0) return COSMA_AMD_OK;
    if (Z > I8_MAX_JOBS) {
        set_error("i8gemm: %d jobs exceed the limit of %d", Z, I8_MAX_JOBS);
        return COSMA_AMD_ERR_UNSUPPORTED;
    }
    auto a16 = [](const void *x) { return (reinterpret_cast<uintptr_t>(x) & 15) == 0; };
    if ((p.lda * BM) >> 31 || (p.ldb * BN) >> 31) {
        set_error("i8gemm: leading dimensions of %lld / %lld bytes are too large", (long long)p.lda, (long long)p.ldb);
        return COSMA_AMD_ERR_UNSUPPORTED;
    }
    if (p.K % (2 * BK) || p.lda % 16 || p.ldb % 16 || p.a_stride % 16 || p.b_stride % 16 || !a16(p.a) || !a16(p.b)) {
        set_error("i8gemm: K must be a multiple of 256 and strides/bases 16-byte aligned");
        return COSMA_AMD_ERR_INVALID_ARG;
    }
    if (p.bf16 && epi != EPI_RES8) {
        set_error("i8gemm: BF16 operands need the residue epilogue");
        return COSMA_AMD_ERR_INVALID_ARG;
    }
    if (epi == EPI_RES8 && (p.ldo % 16 || p.out_stride % 16 || (reinterpret_cast<uintptr_t>(p.out) & 15))) {
        set_error("i8gemm: int8 output leading dimension and stride must be multiples of 16 (16-byte aligned base)");
        return COSMA_AMD_ERR_INVALID_ARG;
    }
    for (int z = 0; z < Z; ++z) {
        if (jobs.npairs[z] < 0 || jobs.npairs[z] > I8_MAX_PAIRS) {
            set_error("i8gemm: job %d has %d pairs", z, jobs.npairs[z]);
            return COSMA_AMD_ERR_INVALID_ARG;
        }
        if (epi == EPI_RES8 && (jobs.modulus[z] < 3 || jobs.modulus[z] > 256 ||
                                (jobs.modulus[z] % 2 == 0 && jobs.modulus[z] != 256))) {
            set_error("i8gemm: modulus %d is neither 256 nor odd in [3, 255]", jobs.modulus[z]);
            return COSMA_AMD_ERR_INVALID_ARG;
        }
    }
    I8Jobs j = jobs;
    int64_t sum_len = 0;   // longest k sum of a job, in elements (bytes)
    for (int z = 0; z < Z; ++z) {
        sum_len = std::max(sum_len, (int64_t)j.npairs[z] * p.K);
        const int m = j.modulus[z] > 0 ? j.modulus[z] : 1;
        j.fmod[z][0] = 1.0f / (float)m;
        j.fmod[z][1] = -(float)m;
        j.fmod[z][2] = (float)(65536 % m);
    }
    const int64_t tiles = ((p.M + BM - 1) / BM) * ((p.N + BN - 1) / BN);
    const dim3 grid((unsigned)tiles, (unsigned)Z), block(NT);
    switch (epi) {
    case EPI_RES8:
        if (p.bf16) hipLaunchKernelGGL((i8gemm_kernel<EPI_RES8, true>), grid, block, 0, s, p, j);
        else if (sum_len <= 32768) hipLaunchKernelGGL((i8gemm_kernel<EPI_RES8, false>), grid, block, 0, s, p, j);
        else hipLaunchKernelGGL((i8gemm_kernel<EPI_RES8, false, true>), grid, block, 0, s, p, j);
        break;
    default: hipLaunchKernelGGL((i8gemm_kernel<EPI_F64ACC, false>), grid, block, 0, s, p, j); break;
    }
    CA_HIP(hipGetLastError());
    return COSMA_AMD_OK;
}

}  // namespace cosma_amd
