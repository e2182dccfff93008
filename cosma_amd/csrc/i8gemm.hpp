// Batched INT8 x INT8 -> INT32 GEMM on the gfx950 matrix cores
// (v_mfma_i32_16x16x64_i8), the engine of the split-precision DGEMM.
//
// Operands are k-contiguous int8 matrices: A is M x K stored as rows of K
// bytes (row i at A + i * lda), B is N x K stored as columns of K bytes
// (column j at B + j * ldb).  All slices (or residues) of one operand live in
// one allocation, slice s at base + s * slice_stride.  The split kernels write
// them in this layout, so every global load is a 16-byte run along k.
//
// One launch computes, for each job z in [0, Z), an M x N result
//     acc_z = sum over the pairs p of job z of  A[pa[z][p]] * B[pb[z][p]]^T
// (a pair list concatenates products along k without moving data), and hands
// acc_z to an epilogue:
//   EPI_RES8    reduce modulo the job's modulus into the symmetric range and
//               store one int8 residue per element (moduli method);
//   EPI_F64ACC  C64 = scale * acc (+ C64 if accumulate) in FP64 (slices method;
//               jobs of one launch must write disjoint outputs).
//
// Tile: 256 x 256 per workgroup of 256 threads (2 x 2 waves of 128 x 128, i.e.
// 8 x 8 MFMA blocks of 16 x 16 per wave, 256 INT32 accumulators per lane in
// AGPRs), k-step 128 (two MFMA depths: every row of a global load is one
// full 128-byte cache line).  LDS tiles use unpadded 128-byte rows with a
// chunk swizzle (chunk ^ (row & 7)) for conflict-free stores and fragment
// reads; two LDS stages (128 KB) double-buffer the steps.
// MFMA i32 16x16x64 lane layout (checked on gfx950): lane l = 16g + i supplies
// A row i / B column i, bytes k = 16g .. 16g+15; it receives D rows 4g..4g+3 of
// column i.
//
// BF16 operands (I8Problem::bf16, EPI_RES8 only): the same byte layout holds
// BF16 numbers with integer values in [-128, 127] (K, lda, ldb and the slice
// strides still count bytes, two per element), and the same 16-byte
// fragments feed v_mfma_f32_16x16x32_bf16.  Products of such integers are
// below 2^14 and FP32 holds every integer below 2^24 exactly, so the FP32
// accumulators stay exact as long as they are reduced modulo the job's
// modulus before they can reach 2^24: every BF16_FLUSH_STEPS k-steps (896
// elements: 896 * 2^14 plus a reduced value below 384 is under 2^24).
#pragma once
#include "common.hpp"

namespace cosma_amd {

constexpr int I8_MAX_PAIRS = 8;
constexpr int I8_MAX_JOBS = 24;

enum I8Epilogue { EPI_RES8 = 1, EPI_F64ACC = 2 };

struct I8Jobs {
    int npairs[I8_MAX_JOBS];
    unsigned char pa[I8_MAX_JOBS][I8_MAX_PAIRS];
    unsigned char pb[I8_MAX_JOBS][I8_MAX_PAIRS];
    int modulus[I8_MAX_JOBS];
    double scale[I8_MAX_JOBS];
    // set by launch_i8gemm from modulus: 1 / m (correctly rounded), -m and
    // 2^16 mod m as floats (kernel arguments, so they stay in scalar registers)
    float fmod[I8_MAX_JOBS][3];
};

struct I8Problem {
    int64_t M, N, K;
    const int8_t *a;      // A slice 0
    int64_t lda;          // bytes between rows of A
    int64_t a_stride;     // bytes between slices of A
    const int8_t *b;
    int64_t ldb;
    int64_t b_stride;
    void *out;            // output of job 0 (int32 / int8 / double, column-major)
    int64_t ldo;          // leading dimension of the outputs (elements)
    int64_t out_stride;   // elements between the outputs of consecutive jobs
    int accumulate;       // EPI_F64ACC: add into the output
    int bf16;             // operands are BF16 integers (EPI_RES8, one pair per job)
};

constexpr int BF16_FLUSH_STEPS = 14;

// Requirements: K % 256 == 0 (bytes); lda, ldb, a_stride, b_stride multiples of 16 and
// 16-byte aligned bases; int8 outputs with ldo and out_stride multiples of 16
// and a 16-byte aligned base (all checked).  Not checkable, hence the callers' duty: every A slice
// holds ceil(M/256)*256 rows and every B slice ceil(N/256)*256 columns (the
// split kernels pad with zeros), so tile loads never need bounds checks.
int launch_i8gemm(const I8Problem &p, const I8Jobs &jobs, int Z, I8Epilogue epi, hipStream_t s);

// Symmetric residue of x modulo m, in [-(m/2), (m-1)/2].  The quotient comes
// from an FP32 estimate: for |x| < 2^31 and m >= 2 the estimate is within one
// of the rounded quotient, so one correction on each side is exact.
__device__ __forceinline__ int sym_mod(int x, int m, float inv_m) {
    const int q = (int)rintf((float)x * inv_m);
    int r = x - q * m;
    r -= (r > (m - 1) / 2) ? m : 0;
    r += (r < -(m / 2)) ? m : 0;
    return r;
}

// The same residue for odd m and |x| < 2^29 in six full-rate FP32
// instructions (sym_mod's 32-bit multiply runs at a quarter of the rate); the
// returned word's low byte is sym_mod(x, m) as a two's-complement byte.
// x = h 2^16 + l (h = x >> 16, l = x & 0xffff, both exact in FP32) and
// c16 = 2^16 mod m give y = h c16 + l, congruent to x, with
// |y| <= 2^13 * 254 + 65535 < 2^21.1: the FMA forming it is exact.
// q = rint(y * fl(1/m)) is the integer nearest to y / m: the product's error
// is below |y| / m * 2^-23 < 0.26 / m, and for odd m the quotient y / m lies
// at least 1 / (2m) from every half-integer.  Hence y - q m is the symmetric
// residue, |r| <= (m - 1) / 2, and the FMA that forms it from y + 1.5 * 2^23
// (exact: the sum lies in [2^23, 2^24), where FP32 holds every integer)
// leaves r + 2^22 in the mantissa, whose low byte is r mod 256.  m = 256
// needs no arithmetic: the residue is x's low byte.
constexpr float SYM_MOD_MAGIC = 12582912.0f;   // 1.5 * 2^23
__device__ __forceinline__ uint32_t sym_mod_odd_bits(int x, float c16, float inv_m, float neg_m) {
    const float yf = fmaf((float)(x >> 16), c16, (float)(x & 0xffff));
    const float q = rintf(yf * inv_m);
    return __float_as_uint(fmaf(q, neg_m, yf + SYM_MOD_MAGIC));
}
// Any int x: x - (q m) 2^16, with q the integer nearest to h / m (the same
// argument; |h| <= 2^15), is congruent to x and below 2^23 in magnitude.
__device__ __forceinline__ uint32_t sym_mod_odd_bits_wide(int x, float c16, float inv_m, float neg_m) {
    const float qm = rintf((float)(x >> 16) * inv_m) * neg_m;   // exact: |q m| < 2^16
    return sym_mod_odd_bits(x + ((int)qm << 16), c16, inv_m, neg_m);
}

}  // namespace cosma_amd
