// Host-side packing of the host-streaming GEMM (see host_copy.hpp).
#include "host_copy.hpp"

#include <algorithm>
#include <cstring>
#include <thread>
#include <vector>

namespace cosma_amd {

void host_copy_2d(double *dst, int64_t dld, const double *src, int64_t sld, int64_t rows, int64_t cols) {
    if (rows <= 0 || cols <= 0) return;
    const size_t bytes = (size_t)rows * (size_t)cols * 8;
    int nt = (int)std::min<size_t>(8, std::max<size_t>(1, bytes >> 22));   // one thread per 4 MB
    nt = (int)std::min<int64_t>(nt, cols);
    auto work = [=](int64_t c0, int64_t c1) {
        for (int64_t c = c0; c < c1; ++c) std::memcpy(dst + c * dld, src + c * sld, (size_t)rows * 8);
    };
    if (nt <= 1) {
        work(0, cols);
        return;
    }
    std::vector<std::thread> th;
    const int64_t per = (cols + nt - 1) / nt;
    for (int t = 0; t < nt; ++t) {
        const int64_t c0 = t * per, c1 = std::min(cols, c0 + per);
        if (c0 < c1) th.emplace_back(work, c0, c1);
    }
    for (auto &x : th) x.join();
}

}  // namespace cosma_amd
