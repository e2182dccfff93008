// Seeded synthetic inputs generated on the device.
//
// The reference's test generator (src/cosma/random_generator.hpp:26-32) seeds
// std::mt19937 from std::random_device, so its inputs cannot be replayed.  This
// fill is counter-based: element i of a stream with seed s is
//   splitmix64(s * 0x9E3779B97F4A7C15 + offset + i)
// mapped to uniform [-1, 1) as ((x >> 11) * 2^-53) * 2 - 1, which is exact in
// FP64, so a CPU generator with the same formula reproduces every value.
// Matrices of several GB are generated in place in HBM instead of being
// produced on the host and copied over PCIe.
#include "common.hpp"

namespace cosma_amd {
namespace {

__device__ __forceinline__ uint64_t splitmix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

// Column-major rows x cols matrix with leading dimension ld; element (r, c)
// takes counter index r + c * rows (the dense column-major position), so the
// values do not depend on ld.
__global__ void fill_uniform_kernel(double *__restrict__ x, int64_t rows, int64_t cols, int64_t ld,
                                    uint64_t base) {
    const int64_t total = rows * cols;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t c = i / rows, r = i - c * rows;
        const uint64_t v = splitmix64(base + (uint64_t)i);
        x[r + c * ld] = (double)(v >> 11) * 0x1.0p-53 * 2.0 - 1.0;
    }
}

// a = (u - 0.5) * exp(phi * z), u ~ U[0,1), z ~ N(0,1) by Box-Muller; element
// i uses stream positions 3i, 3i+1, 3i+2.  Values depend on the device's
// log/cos/exp, so they are reproducible on the device, not bit-for-bit on
// the host.
__global__ void fill_lognormal_kernel(double *__restrict__ x, int64_t rows, int64_t cols, int64_t ld,
                                      uint64_t base, double phi) {
    const int64_t total = rows * cols;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t c = i / rows, r = i - c * rows;
        const uint64_t c3 = base + 3ull * (uint64_t)i;
        const double u = (double)(splitmix64(c3) >> 11) * 0x1.0p-53;
        const double u1 = (double)((splitmix64(c3 + 1) >> 11) + 1) * 0x1.0p-53;   // (0, 1]
        const double u2 = (double)(splitmix64(c3 + 2) >> 11) * 0x1.0p-53;
        const double z = sqrt(-2.0 * log(u1)) * cospi(2.0 * u2);
        x[r + c * ld] = (u - 0.5) * exp(phi * z);
    }
}

}  // namespace

int launch_fill_lognormal(double *x, int64_t rows, int64_t cols, int64_t ld, uint64_t seed, double phi,
                          hipStream_t s) {
    if (rows == 0 || cols == 0) return COSMA_AMD_OK;
    int64_t blocks = (rows * cols + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    const uint64_t base = seed * 0x9E3779B97F4A7C15ULL;
    hipLaunchKernelGGL(fill_lognormal_kernel, dim3((unsigned)blocks), dim3(256), 0, s, x, rows, cols, ld, base, phi);
    CA_HIP(hipGetLastError());
    return COSMA_AMD_OK;
}

int launch_fill_uniform(double *x, int64_t rows, int64_t cols, int64_t ld, uint64_t seed,
                        uint64_t offset, hipStream_t s) {
    if (rows == 0 || cols == 0) return COSMA_AMD_OK;
    const int64_t total = rows * cols;
    int64_t blocks = (total + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    const uint64_t base = seed * 0x9E3779B97F4A7C15ULL + offset;
    hipLaunchKernelGGL(fill_uniform_kernel, dim3((unsigned)blocks), dim3(256), 0, s, x, rows, cols, ld, base);
    CA_HIP(hipGetLastError());
    return COSMA_AMD_OK;
}

}  // namespace cosma_amd
