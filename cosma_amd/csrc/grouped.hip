// Grouped (batched) DGEMM plans: many independent problems in one persistent
// launch, replayable as a HIP graph, with a built-in autotuner.
//
// No counterpart in the reference, whose local multiplies are one GEMM each
// (src/cosma/local_multiply.cpp:299-381); on one rank the user-specified
// sequential steps (src/cosma/multiply.cpp:460-588) and any batch of small
// multiplies are the natural clients.
//
// A plan copies the problem descriptors to the device once, sorted by
// decreasing cost (k * m * n) so that the static round-robin assignment of
// tiles to persistent workgroups starts with the longest tiles (a
// longest-processing-time order).  Running a plan is a single kernel launch
// with no allocation, no memset and no host synchronisation, so it can be
// captured; cosma_amd_grouped_graph_run captures it once into a hipGraph and
// replays that.  The tuner times every tile configuration (LDS-staged 64x64,
// 128x64, 128x128; direct 64x64 and 32x32) with persistent grids of 1, 2, 4
// and 8 workgroups per CU and one workgroup per tile, keeps the fastest, and
// remembers it process-wide for plans with the same problem shapes.
#include <algorithm>
#include <map>
#include <mutex>
#include <numeric>
#include <vector>

#include "common.hpp"

struct cosma_amd_grouped_s {
    cosma_amd_handle_t h = nullptr;
    int nprob = 0;
    uint64_t signature = 0;
    // per tile configuration 1..5; configuration 4 (64 x 64, direct) shares
    // configuration 1's descriptors
    cosma_amd::GroupedProblem *dev[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    int64_t total_tiles[6] = {0, 0, 0, 0, 0, 0};
    int cfg = 1;
    int grid = 1;
    hipStream_t capture = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace cosma_amd {
namespace {

std::mutex g_tune_mutex;
std::map<uint64_t, std::pair<int, int>> g_tuned;   // shape signature -> (tile config, grid)

uint64_t fnv(uint64_t h, uint64_t v) {
    for (int i = 0; i < 8; ++i) {
        h ^= (v >> (8 * i)) & 0xff;
        h *= 0x100000001b3ULL;
    }
    return h;
}

void drop_graph(cosma_amd_grouped_t p) {
    if (p->exec) hipGraphExecDestroy(p->exec);
    if (p->graph) hipGraphDestroy(p->graph);
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

int default_grid(cosma_amd_grouped_t p, int cfg) {
    // persistent: enough workgroups to fill every CU several times over
    const int64_t g = std::min<int64_t>(p->total_tiles[cfg], (int64_t)p->h->n_cu * 4);
    return (int)std::max<int64_t>(1, g);
}

}  // namespace

int grouped_create(cosma_amd_handle_t h, const cosma_amd_gemm_desc *d, int64_t count, int tile_config,
                   cosma_amd_grouped_t *out) {
    *out = nullptr;
    if (count < 0 || count > 0x7fffffff || (count > 0 && !d)) {
        set_error("grouped dgemm: invalid problem list");
        return COSMA_AMD_ERR_INVALID_ARG;
    }
    if (tile_config < 0 || tile_config > 5) {
        set_error("grouped dgemm: tile_config %d is not 0..5", tile_config);
        return COSMA_AMD_ERR_INVALID_ARG;
    }
    std::vector<GroupedProblem> probs;
    std::vector<double> cost;
    uint64_t sig = 0xcbf29ce484222325ULL;
    for (int64_t i = 0; i < count; ++i) {
        const cosma_amd_gemm_desc &q = d[i];
        const char ta = norm_trans(q.transa), tb = norm_trans(q.transb);
        if (!ta || !tb || q.m < 0 || q.n < 0 || q.k < 0) {
            set_error("grouped dgemm: problem %lld has an invalid transpose or a negative dimension", (long long)i);
            return COSMA_AMD_ERR_INVALID_ARG;
        }
        if (q.m == 0 || q.n == 0) continue;                 // nothing to do
        const bool scale_only = (q.k == 0 || q.alpha == 0.0);
        const int64_t ar = ta == 'N' ? q.m : q.k, br = tb == 'N' ? q.k : q.n;
        if (q.ldc < std::max<int64_t>(1, q.m) || !q.c || !on_device(q.c) ||
            (!scale_only && (q.lda < std::max<int64_t>(1, ar) || q.ldb < std::max<int64_t>(1, br) || !q.a || !q.b ||
                             !on_device(q.a) || !on_device(q.b)))) {
            set_error("grouped dgemm: problem %lld has a leading dimension too small or an operand that is not "
                      "device memory", (long long)i);
            return COSMA_AMD_ERR_INVALID_ARG;
        }
        GroupedProblem g{};
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
        grouped_problem_flags(ta, tb, q.m, q.n, q.a, q.lda, q.b, q.ldb, q.c, q.ldc, &g.flags, &g.vec_c);
        probs.push_back(g);
        cost.push_back(scale_only ? 0.0 : (double)q.m * q.n * q.k);
        sig = fnv(fnv(fnv(fnv(sig, q.m), q.n), scale_only ? 0 : q.k), g.flags);
    }
    std::vector<int> order(probs.size());
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return cost[x] > cost[y]; });

    auto p = new cosma_amd_grouped_s();
    p->h = h;
    p->nprob = (int)probs.size();
    p->signature = fnv(sig, (uint64_t)probs.size());
    int rc = COSMA_AMD_OK;
    for (int cfg : {1, 2, 3, 5}) {
        if (rc != COSMA_AMD_OK) break;
        int bm, bn;
        grouped_tile_shape(cfg, &bm, &bn);
        std::vector<GroupedProblem> sorted;
        sorted.reserve(probs.size());
        int64_t t = 0;
        for (int i : order) {
            GroupedProblem g = probs[i];
            g.tile_start = t;
            t += ((g.m + bm - 1) / bm) * ((g.n + bn - 1) / bn);
            sorted.push_back(g);
        }
        p->total_tiles[cfg] = t;
        if (sorted.empty()) continue;
        // descriptors, then their first tile indices (the kernels' search array)
        std::vector<int64_t> starts;
        for (const GroupedProblem &g : sorted) starts.push_back(g.tile_start);
        const size_t dbytes = sorted.size() * sizeof(GroupedProblem), sbytes = starts.size() * sizeof(int64_t);
        hipError_t e = dev_alloc(reinterpret_cast<void **>(&p->dev[cfg]), dbytes + sbytes);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            set_error("grouped dgemm: descriptor allocation failed (%s)", hipGetErrorString(e));
            rc = COSMA_AMD_ERR_ALLOC;
            break;
        }
        e = hipMemcpy(p->dev[cfg], sorted.data(), dbytes, hipMemcpyHostToDevice);
        if (e == hipSuccess)
            e = hipMemcpy(reinterpret_cast<char *>(p->dev[cfg]) + dbytes, starts.data(), sbytes, hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            set_error("grouped dgemm: descriptor upload failed (%s)", hipGetErrorString(e));
            rc = COSMA_AMD_ERR_HIP;
        }
    }
    if (rc != COSMA_AMD_OK) {
        grouped_destroy(p);
        return rc;
    }
    p->dev[4] = p->dev[1];
    p->total_tiles[4] = p->total_tiles[1];
    {
        std::lock_guard<std::mutex> lock(g_tune_mutex);
        auto it = g_tuned.find(p->signature);
        if (tile_config == 0 && it != g_tuned.end()) {
            p->cfg = it->second.first;
            p->grid = it->second.second;
        } else {
            p->cfg = tile_config == 0 ? 1 : tile_config;
            p->grid = default_grid(p, p->cfg);
        }
    }
    *out = p;
    return COSMA_AMD_OK;
}

int grouped_run(cosma_amd_grouped_t p, hipStream_t s) {
    CA_HIP(hipSetDevice(p->h->device));
    return launch_grouped(p->cfg, p->dev[p->cfg], p->nprob, p->total_tiles[p->cfg], p->grid, s);
}

int grouped_graph_run(cosma_amd_grouped_t p, hipStream_t s) {
    if (p->nprob == 0) return COSMA_AMD_OK;
    CA_HIP(hipSetDevice(p->h->device));
    if (!p->exec) {
        if (!p->capture) CA_HIP(hipStreamCreateWithFlags(&p->capture, hipStreamNonBlocking));
        CA_HIP(hipStreamBeginCapture(p->capture, hipStreamCaptureModeThreadLocal));
        const int rc = grouped_run(p, p->capture);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(p->capture, &g);
        if (rc != COSMA_AMD_OK) {
            if (g) hipGraphDestroy(g);
            return rc;
        }
        CA_HIP(e);
        p->graph = g;
        CA_HIP(hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0));
    }
    CA_HIP(hipGraphLaunch(p->exec, s));
    return COSMA_AMD_OK;
}

int grouped_tune(cosma_amd_grouped_t p, hipStream_t s) {
    if (p->nprob == 0) return COSMA_AMD_OK;
    CA_HIP(hipSetDevice(p->h->device));
    hipEvent_t a, b;
    CA_HIP(hipEventCreate(&a));
    CA_HIP(hipEventCreate(&b));
    float best = -1.0f;
    int best_cfg = p->cfg, best_grid = p->grid;
    int rc = COSMA_AMD_OK;
    for (int cfg = 1; cfg <= 5 && rc == COSMA_AMD_OK; ++cfg) {
        const int64_t tiles = p->total_tiles[cfg];
        std::vector<int64_t> grids;
        for (int per_cu : {1, 2, 4, 8}) grids.push_back(std::min<int64_t>(tiles, (int64_t)p->h->n_cu * per_cu));
        grids.push_back(std::min<int64_t>(tiles, 0x7fffffff));
        std::sort(grids.begin(), grids.end());
        grids.erase(std::unique(grids.begin(), grids.end()), grids.end());
        for (int64_t g : grids) {
            if (g < 1) continue;
            if ((rc = launch_grouped(cfg, p->dev[cfg], p->nprob, tiles, (int)g, s))) break;   // warm-up
            CA_HIP(hipEventRecord(a, s));
            for (int r = 0; r < 3 && rc == COSMA_AMD_OK; ++r) rc = launch_grouped(cfg, p->dev[cfg], p->nprob, tiles, (int)g, s);
            CA_HIP(hipEventRecord(b, s));
            CA_HIP(hipEventSynchronize(b));
            float ms = 0.0f;
            CA_HIP(hipEventElapsedTime(&ms, a, b));
            if (best < 0.0f || ms < best) {
                best = ms;
                best_cfg = cfg;
                best_grid = (int)g;
            }
        }
    }
    hipEventDestroy(a);
    hipEventDestroy(b);
    if (rc != COSMA_AMD_OK) return rc;
    if (best_cfg != p->cfg || best_grid != p->grid) drop_graph(p);
    p->cfg = best_cfg;
    p->grid = best_grid;
    std::lock_guard<std::mutex> lock(g_tune_mutex);
    g_tuned[p->signature] = {best_cfg, best_grid};
    return COSMA_AMD_OK;
}

int grouped_set_config(cosma_amd_grouped_t p, int tile_config, int grid) {
    if (tile_config < 1 || tile_config > 5) {
        set_error("grouped dgemm: tile_config %d is not 1..5", tile_config);
        return COSMA_AMD_ERR_INVALID_ARG;
    }
    const int g = grid > 0 ? (int)std::min<int64_t>(grid, std::max<int64_t>(1, p->total_tiles[tile_config]))
                           : default_grid(p, tile_config);
    if (tile_config != p->cfg || g != p->grid) drop_graph(p);
    p->cfg = tile_config;
    p->grid = g;
    return COSMA_AMD_OK;
}

void grouped_info(cosma_amd_grouped_t p, int *tile_config, int *grid, int64_t *tiles) {
    if (tile_config) *tile_config = p->cfg;
    if (grid) *grid = p->grid;
    if (tiles) *tiles = p->total_tiles[p->cfg];
}

void grouped_destroy(cosma_amd_grouped_t p) {
    if (!p) return;
    hipSetDevice(p->h->device);
    drop_graph(p);
    if (p->capture) hipStreamDestroy(p->capture);
    for (int cfg : {1, 2, 3, 5})   // dev[4] aliases dev[1]
        if (p->dev[cfg]) hipFree(p->dev[cfg]);
    delete p;
}

}  // namespace cosma_amd
