// Local block transform: dst <- alpha * op(src) + beta * dst (column-major).
//
// Replaces COSTA's local block copy/transpose (costa::memory::copy_and_transpose,
// call site benchmarks/transpose.cpp:46-47) and the per-block alpha/beta update
// that costa::transformer applies when COSMA reshuffles operands between the
// ScaLAPACK layout and its own (src/cosma/cosma_pxgemm.cpp:338-364,
// src/cosma/multiply.cpp:186-204).  On one rank a ScaLAPACK block-cyclic local
// array is the global matrix in column-major order with leading dimension lld
// (numroc with one process returns the full extent, src/cosma/scalapack.cpp:49-87),
// so the reorder reduces to this strided copy / transpose.
//
// op = 'N': each thread streams 2 consecutive doubles of a column (16-byte
//           accesses when both leading dimensions and pointers allow it).
// op = 'T': 64 x 64 tiles staged through LDS with a one-double row pad, so
//           both the read of src and the write of dst walk memory contiguously
//           (rows of the tile are read along src columns and written along
//           dst columns); the pad spreads the column reads over all banks.
//           With aligned pointers and even leading dimensions the global
//           accesses are 16 bytes wide (two consecutive elements) and
//           non-temporal (4.87 -> 5.2 TB/s in the probe of the variants), and
//           a workgroup takes several tiles of a column, loading the next
//           one while it writes the current one (16384^2: 0.865 -> 0.835 ms).
// alpha == 1 and beta == 0 copies values bit-for-bit (no arithmetic at all);
// beta == 0 never reads dst.
#include "common.hpp"

namespace cosma_amd {
namespace {

constexpr int TT = 64;       // transpose tile edge
constexpr int TNT = 256;     // threads per transpose workgroup

__device__ __forceinline__ double combine(double x, double alpha, double beta, const double *d, int mode) {
    // mode 0: pure copy; 1: alpha*x (dst not read); 2: alpha*x + beta*dst
    if (mode == 0) return x;
    if (mode == 1) return alpha * x;
    return fma(beta, *d, alpha * x);
}

// dst(rows x cols) <- f(src(rows x cols)).  Grid: x over row-pairs, y over
// pairs of columns: a thread loads its two 16-byte pieces (one per column)
// before it stores either, and the streamed data bypass the caches
// (non-temporal loads and stores): 16384^2 copy 6.13 -> 6.33 TB/s in a probe
// of the variants (one column per thread, 2, 4 or 8 columns, flat 1-D).
__global__ void copy_kernel(int64_t rows, int64_t cols, double alpha,
                            const double *__restrict__ src, int64_t lds, double beta,
                            double *__restrict__ dst, int64_t ldd, int mode, int vec) {
    const int64_t r = 2 * ((int64_t)blockIdx.x * blockDim.x + threadIdx.x);
    if (r >= rows) return;
    for (int64_t c0 = 2 * (int64_t)blockIdx.y; c0 < cols; c0 += 2 * (int64_t)gridDim.y) {
        const int nc = c0 + 1 < cols ? 2 : 1;
        if (vec && r + 1 < rows) {
            f64x2 v[2];
#pragma unroll
            for (int u = 0; u < 2; ++u)
                if (u < nc) v[u] = __builtin_nontemporal_load(reinterpret_cast<const f64x2 *>(src + r + (c0 + u) * lds));
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                if (u >= nc) continue;
                double *d = dst + r + (c0 + u) * ldd;
                f64x2 o;
                o.x = combine(v[u].x, alpha, beta, d, mode);
                o.y = combine(v[u].y, alpha, beta, d + 1, mode);
                __builtin_nontemporal_store(o, reinterpret_cast<f64x2 *>(d));
            }
        } else {
            for (int u = 0; u < nc; ++u) {
                const double *s = src + r + (c0 + u) * lds;
                double *d = dst + r + (c0 + u) * ldd;
                d[0] = combine(s[0], alpha, beta, d, mode);
                if (r + 1 < rows) d[1] = combine(s[1], alpha, beta, d + 1, mode);
            }
        }
    }
}

// dst(rows x cols)[i, j] <- f(src[j, i]); src is cols x rows with leading dim lds.
// blockIdx.x -> 64-row block of dst (= 64-column block of src),
// blockIdx.y -> 64-column block of dst (= 64-row block of src).
__global__ __launch_bounds__(TNT) void transpose_kernel(int64_t rows, int64_t cols, double alpha,
                                                        const double *__restrict__ src, int64_t lds,
                                                        double beta, double *__restrict__ dst,
                                                        int64_t ldd, int mode) {
    __shared__ double tile[TT][TT + 1];
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    const int64_t i0 = (int64_t)blockIdx.x * TT;   // dst rows / src cols
    for (int64_t j0 = (int64_t)blockIdx.y * TT; j0 < cols; j0 += (int64_t)gridDim.y * TT) {
        // read: src rows j0+tx (contiguous), src cols i0+ty+4r
#pragma unroll
        for (int r = 0; r < TT / 4; ++r) {
            const int il = ty + 4 * r;
            const int64_t i = i0 + il, j = j0 + tx;
            if (i < rows && j < cols) tile[il][tx] = src[j + i * lds];
        }
        __syncthreads();
        // write: dst rows i0+tx (contiguous), dst cols j0+ty+4r
#pragma unroll
        for (int r = 0; r < TT / 4; ++r) {
            const int jl = ty + 4 * r;
            const int64_t i = i0 + tx, j = j0 + jl;
            if (i < rows && j < cols) {
                double *d = dst + i + j * ldd;
                *d = combine(tile[tx][jl], alpha, beta, d, mode);
            }
        }
        __syncthreads();
    }
}

// Same transpose with 16-byte global accesses: every load takes two
// consecutive src elements of one src column, every store writes two
// consecutive dst elements of one dst column.  Needs 16-byte aligned
// pointers and even leading dimensions; tiles that cross the matrix edge take
// the element-wise path.
__global__ __launch_bounds__(TNT) void transpose_vec_kernel(int64_t rows, int64_t cols, double alpha,
                                                            const double *__restrict__ src, int64_t lds,
                                                            double beta, double *__restrict__ dst,
                                                            int64_t ldd, int mode) {
    __shared__ double tile[TT][TT + 1];
    const int t = threadIdx.x;
    const int64_t i0 = (int64_t)blockIdx.x * TT;   // dst rows / src cols
    const int64_t jstep = (int64_t)gridDim.y * TT;
    auto full = [&](int64_t j0) { return (i0 + TT <= rows) && (j0 + TT <= cols); };
    // src column i0 + il, rows j0 + c2, c2 + 1
    auto load = [&](int64_t j0, f64x2 (&v)[8]) {
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int idx = t + TNT * q, il = idx >> 5, c2 = (idx & 31) * 2;
            v[q] = __builtin_nontemporal_load(reinterpret_cast<const f64x2 *>(src + (j0 + c2) + (i0 + il) * lds));
        }
    };
    // A workgroup walks the tiles j0, j0 + jstep, ... of its column of
    // tiles; the loads of the next full tile are issued right after the
    // current one is in LDS, so they are in flight during its stores.
    f64x2 v[8];
    int64_t j0 = (int64_t)blockIdx.y * TT;
    if (j0 < cols && full(j0)) load(j0, v);
    for (; j0 < cols; j0 += jstep) {
        if (full(j0)) {
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const int idx = t + TNT * q, il = idx >> 5, c2 = (idx & 31) * 2;
                tile[il][c2] = v[q].x;
                tile[il][c2 + 1] = v[q].y;
            }
            __syncthreads();
            if (j0 + jstep < cols && full(j0 + jstep)) load(j0 + jstep, v);
#pragma unroll
            for (int q = 0; q < 8; ++q) {     // dst column j0 + jl, rows i0 + r2, r2 + 1
                const int idx = t + TNT * q, jl = idx >> 5, r2 = (idx & 31) * 2;
                f64x2 *d = reinterpret_cast<f64x2 *>(dst + (i0 + r2) + (j0 + jl) * ldd);
                f64x2 o;
                if (mode == 0) {
                    o.x = tile[r2][jl];
                    o.y = tile[r2 + 1][jl];
                } else if (mode == 1) {
                    o.x = alpha * tile[r2][jl];
                    o.y = alpha * tile[r2 + 1][jl];
                } else {
                    const f64x2 old = *d;
                    o.x = fma(beta, old.x, alpha * tile[r2][jl]);
                    o.y = fma(beta, old.y, alpha * tile[r2 + 1][jl]);
                }
                __builtin_nontemporal_store(o, d);
            }
        } else {
            const int tx = t & 63, ty = t >> 6;
#pragma unroll
            for (int r = 0; r < TT / 4; ++r) {
                const int il = ty + 4 * r;
                const int64_t i = i0 + il, j = j0 + tx;
                if (i < rows && j < cols) tile[il][tx] = src[j + i * lds];
            }
            __syncthreads();
            if (j0 + jstep < cols && full(j0 + jstep)) load(j0 + jstep, v);
#pragma unroll
            for (int r = 0; r < TT / 4; ++r) {
                const int jl = ty + 4 * r;
                const int64_t i = i0 + tx, j = j0 + jl;
                if (i < rows && j < cols) {
                    double *d = dst + i + j * ldd;
                    *d = combine(tile[tx][jl], alpha, beta, d, mode);
                }
            }
        }
        __syncthreads();
    }
}

inline bool aligned16(const void *p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; }

}  // namespace

int launch_block_transform(char op, int64_t rows, int64_t cols, double alpha,
                           const double *src, int64_t lds, double beta,
                           double *dst, int64_t ldd, hipStream_t s) {
    if (rows == 0 || cols == 0) return COSMA_AMD_OK;
    const int mode = (alpha == 1.0 && beta == 0.0) ? 0 : (beta == 0.0 ? 1 : 2);
    if (op == 'N') {
        const int vec = aligned16(src) && aligned16(dst) && (lds % 2 == 0) && (ldd % 2 == 0);
        const int64_t pairs = (rows + 1) / 2;
        const int64_t bx = (pairs + 255) / 256;
        const int64_t by = std::min<int64_t>((cols + 1) / 2, 65535);
        if (bx > 0x7fffffffLL) {
            set_error("block_transform: %lld rows exceed the grid limit", (long long)rows);
            return COSMA_AMD_ERR_UNSUPPORTED;
        }
        hipLaunchKernelGGL(copy_kernel, dim3((unsigned)bx, (unsigned)by), dim3(256), 0, s,
                           rows, cols, alpha, src, lds, beta, dst, ldd, mode, vec);
    } else {
        const int64_t bx = (rows + TT - 1) / TT;
        int64_t by = (cols + TT - 1) / TT;
        if (by > 65535) by = 65535;
        if (bx > 0x7fffffffLL) {
            set_error("block_transform: %lld rows exceed the grid limit", (long long)rows);
            return COSMA_AMD_ERR_UNSUPPORTED;
        }
        const bool vec = aligned16(src) && aligned16(dst) && (lds % 2 == 0) && (ldd % 2 == 0);
        // with plenty of workgroups, four tiles each (the vectorised kernel
        // prefetches the next tile of a workgroup during the current one)
        if (vec && bx * by > 4096) by = (by + 3) / 4;
        if (vec)
            hipLaunchKernelGGL(transpose_vec_kernel, dim3((unsigned)bx, (unsigned)by), dim3(TNT), 0, s,
                               rows, cols, alpha, src, lds, beta, dst, ldd, mode);
        else
            hipLaunchKernelGGL(transpose_kernel, dim3((unsigned)bx, (unsigned)by), dim3(TNT), 0, s,
                               rows, cols, alpha, src, lds, beta, dst, ldd, mode);
    }
    CA_HIP(hipGetLastError());
    return COSMA_AMD_OK;
}

}  // namespace cosma_amd
