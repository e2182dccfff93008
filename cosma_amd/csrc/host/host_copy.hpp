// Host-side packing of the host-streaming GEMM (plain C++, no HIP): the
// copies between pageable caller memory and the pinned staging slots.
#pragma once
#include <cstdint>

namespace cosma_amd {

// Column-major 2D copy of a rows x cols block between host buffers (leading
// dimensions dld, sld, in elements), split over up to 8 threads by columns
// (one thread per 4 MB).
void host_copy_2d(double *dst, int64_t dld, const double *src, int64_t sld, int64_t rows, int64_t cols);

}  // namespace cosma_amd
