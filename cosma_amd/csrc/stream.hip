// Host-tile streaming DGEMM: C <- alpha * op(A) * op(B) + beta * C where any
// of A, B, C may live in host memory (pinned or pageable) or on the device.
//
// Replaces Tiled-MM's gpu::gemm(mm_handle&, ta, tb, m, n, k, alpha, a, lda, b,
// ldb, beta, c, ldc, pin_host_buffers, copy_c_back), which COSMA calls on
// host-resident operands (src/cosma/local_multiply.cpp:245-260) after pinning
// them itself (src/cosma/local_multiply.cpp:341-349).
//
// Schedule.  C is cut into tile_m x tile_n tiles and k into tile_k panels
// (the COSMA_GPU_MAX_TILE_{M,N,K} sizes of the handle).  Three kinds of work
// run on separate streams and are ordered only by events:
//   * one host->device stream uploads operand blocks, in the order the tiles
//     need them;
//   * compute streams run the FP64 MFMA DGEMM of each (tile, panel) as soon
//     as that block's upload event has fired; the products of one tile
//     accumulate in order on one stream (the first with the tile's starting
//     beta, the rest with 1);
//   * one device->host stream returns each finished C tile while later tiles
//     are still being uploaded and multiplied (PCIe is full duplex).
// Resident mode (the default when the operands fit in the device-memory
// budget — 288 GB of HBM makes this the common case): every block of A and B
// is uploaded once into a device copy of the whole operand and reused by all
// tiles that need it, so the host link carries A, B, (C) once instead of
// re-streaming panels per tile.  Tiles are visited in growing squares; the
// first half of the panels runs panel-major, the rest tile by tile; a host C
// with beta != 0 is uploaded last per tile and added as beta * C after the
// tile's products, so the link first carries the blocks that start compute;
// all products run on one compute stream in upload order.
// Bounded mode (operands larger than the budget): Tiled-MM's scheme — per
// tile, panels pass through two device slots per operand whose reuse waits
// on the compute that last read them, panel 0 applying beta to the C tile,
// with the tiles round-robin over the COSMA_GPU_STREAMS compute streams.
//
// Host memory.  Pinned host buffers (hipHostMalloc) are copied directly by DMA.
// Pageable buffers go through a ring of pinned staging slots: the host packs
// a block into a slot (several threads for large blocks) and the slot is
// DMA'd while the next one is packed; returned C tiles are unpacked from
// their slots once their copy event has fired.  Registering caller memory in
// place (hipHostRegister, the reference's approach) is not used.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>

#include "common.hpp"
#include "host/host_copy.hpp"

namespace cosma_amd {
namespace {

enum Kind { DEVICE, PINNED, PAGEABLE };

// COSMA_AMD_TRACE=1 prints the schedule of each streaming call to stderr.
bool trace_on() {
    static const bool on = [] {
        const char *v = std::getenv("COSMA_AMD_TRACE");
        return v && v[0] == '1';
    }();
    return on;
}
#define TRACE(...)                                      \
    do {                                                \
        if (trace_on()) {                               \
            std::fprintf(stderr, "[cosma_amd] " __VA_ARGS__); \
            std::fputc('\n', stderr);                   \
        }                                               \
    } while (0)

Kind classify(const void *p) {
    hipPointerAttribute_t attr;
    hipError_t e = hipPointerGetAttributes(&attr, p);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return PAGEABLE;
    }
    if (attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged) return DEVICE;
    if (attr.type == hipMemoryTypeHost) return PINNED;
    return PAGEABLE;
}

constexpr size_t SLOT_BYTES = 64ull << 20;   // pinned staging slot
constexpr int N_SLOTS = 4;                   // per direction

int grow(DevBuf &b, size_t bytes) {
    if (b.bytes >= bytes) return COSMA_AMD_OK;
    if (b.ptr) {
        CA_HIP(hipFree(b.ptr));
        b.ptr = nullptr;
        b.bytes = 0;
    }
    hipError_t e = dev_alloc(&b.ptr, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        b.ptr = nullptr;
        set_error("host streaming: hipMalloc of %zu bytes failed (%s)", bytes, hipGetErrorString(e));
        return COSMA_AMD_ERR_ALLOC;
    }
    b.bytes = bytes;
    return COSMA_AMD_OK;
}

int get_event(cosma_amd_handle_t h, hipEvent_t *ev) {
    if (h->events_used == h->events.size()) {
        hipEvent_t e;
        CA_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        h->events.push_back(e);
    }
    *ev = h->events[h->events_used++];
    return COSMA_AMD_OK;
}

int record(cosma_amd_handle_t h, hipStream_t s, hipEvent_t *ev) {
    int rc = get_event(h, ev);
    if (rc) return rc;
    CA_HIP(hipEventRecord(*ev, s));
    return COSMA_AMD_OK;
}

int ensure_slots(std::vector<StagingSlot> &slots) {
    if (!slots.empty()) return COSMA_AMD_OK;
    // all N_SLOTS or none: a failure releases the slots made so far, so a
    // later call starts over with a full ring
    auto release = [&]() {
        for (auto &s : slots) {
            if (s.ev) (void)hipEventDestroy(s.ev);
            if (s.ptr) (void)hipHostFree(s.ptr);
        }
        slots.clear();
    };
    for (int i = 0; i < N_SLOTS; ++i) {
        StagingSlot s;
        hipError_t e = pinned_alloc(&s.ptr, SLOT_BYTES);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            release();
            set_error("host streaming: pinned staging allocation failed (%s)", hipGetErrorString(e));
            return COSMA_AMD_ERR_ALLOC;
        }
        s.bytes = SLOT_BYTES;
        e = hipEventCreateWithFlags(&s.ev, hipEventDisableTiming);
        if (e != hipSuccess) {
            (void)hipHostFree(s.ptr);
            release();
            return hip_fail(e, "hipEventCreateWithFlags", __FILE__, __LINE__);
        }
        slots.push_back(s);
    }
    return COSMA_AMD_OK;
}


// Finishes a pending copy-out of a D2H staging slot into pageable memory.
int drain_slot(StagingSlot &s) {
    if (!s.out) return COSMA_AMD_OK;
    CA_HIP(hipEventSynchronize(s.ev));
    host_copy_2d(s.out, s.out_ld, static_cast<const double *>(s.ptr), s.rows, s.rows, s.cols);
    s.out = nullptr;
    return COSMA_AMD_OK;
}

// Chunking of a rows x cols block into pieces that fit one staging slot.
void chunk_shape(int64_t rows, int64_t cols, int64_t *rc, int64_t *cc) {
    const int64_t cap = (int64_t)(SLOT_BYTES / 8);
    *rc = std::min(rows, cap);
    *cc = std::max<int64_t>(1, std::min(cols, cap / *rc));
}

// Host block (rows x cols, leading dim sld) -> device (leading dim dld) on h->h2d.
int upload(cosma_amd_handle_t h, double *dst, int64_t dld, const double *src, int64_t sld,
           int64_t rows, int64_t cols, Kind kind) {
    if (rows == 0 || cols == 0) return COSMA_AMD_OK;
    if (kind == PINNED) {
        CA_HIP(hipMemcpy2DAsync(dst, dld * 8, src, sld * 8, rows * 8, cols, hipMemcpyHostToDevice, h->h2d));
        return COSMA_AMD_OK;
    }
    int rc = ensure_slots(h->h2d_slots);
    if (rc) return rc;
    int64_t rch, cch;
    chunk_shape(rows, cols, &rch, &cch);
    for (int64_t c0 = 0; c0 < cols; c0 += cch) {
        for (int64_t r0 = 0; r0 < rows; r0 += rch) {
            const int64_t r = std::min(rch, rows - r0), c = std::min(cch, cols - c0);
            StagingSlot &s = h->h2d_slots[h->h2d_next++ % h->h2d_slots.size()];
            CA_HIP(hipEventSynchronize(s.ev));   // previous DMA out of this slot is done
            host_copy_2d(static_cast<double *>(s.ptr), r, src + r0 + c0 * sld, sld, r, c);
            CA_HIP(hipMemcpy2DAsync(dst + r0 + c0 * dld, dld * 8, s.ptr, r * 8, r * 8, c,
                                    hipMemcpyHostToDevice, h->h2d));
            CA_HIP(hipEventRecord(s.ev, h->h2d));
        }
    }
    return COSMA_AMD_OK;
}

// Device block -> host on h->d2h (pageable destinations finish in drain_slot).
int download(cosma_amd_handle_t h, double *dst, int64_t dld, const double *src, int64_t sld,
             int64_t rows, int64_t cols, Kind kind) {
    if (rows == 0 || cols == 0) return COSMA_AMD_OK;
    if (kind == PINNED) {
        CA_HIP(hipMemcpy2DAsync(dst, dld * 8, src, sld * 8, rows * 8, cols, hipMemcpyDeviceToHost, h->d2h));
        return COSMA_AMD_OK;
    }
    int rc = ensure_slots(h->d2h_slots);
    if (rc) return rc;
    int64_t rch, cch;
    chunk_shape(rows, cols, &rch, &cch);
    for (int64_t c0 = 0; c0 < cols; c0 += cch) {
        for (int64_t r0 = 0; r0 < rows; r0 += rch) {
            const int64_t r = std::min(rch, rows - r0), c = std::min(cch, cols - c0);
            StagingSlot &s = h->d2h_slots[h->d2h_next++ % h->d2h_slots.size()];
            rc = drain_slot(s);
            if (rc) return rc;
            CA_HIP(hipMemcpy2DAsync(s.ptr, r * 8, src + r0 + c0 * sld, sld * 8, r * 8, c,
                                    hipMemcpyDeviceToHost, h->d2h));
            CA_HIP(hipEventRecord(s.ev, h->d2h));
            s.out = dst + r0 + c0 * dld;
            s.out_ld = dld;
            s.rows = r;
            s.cols = c;
        }
    }
    return COSMA_AMD_OK;
}

// After a failed call: waits for the handle's streams (ignoring their
// errors) and drops every pending copy-out of the D2H staging slots, so that
// no later call copies a stale tile into a caller buffer of this one.
void abandon(cosma_amd_handle_t h) {
    (void)hipStreamSynchronize(h->h2d);
    for (auto s : h->streams) (void)hipStreamSynchronize(s);
    (void)hipStreamSynchronize(h->d2h);
    (void)hipGetLastError();
    for (auto &s : h->d2h_slots) s.out = nullptr;
}

int sync_all(cosma_amd_handle_t h) {
    CA_HIP(hipStreamSynchronize(h->h2d));
    for (auto s : h->streams) CA_HIP(hipStreamSynchronize(s));
    CA_HIP(hipStreamSynchronize(h->d2h));
    for (auto &s : h->d2h_slots) {
        int rc = drain_slot(s);
        if (rc) return rc;
    }
    return COSMA_AMD_OK;
}

// Position of block (row0, col0) of op(X) inside the stored (column-major) X.
inline const double *op_block(const double *x, int64_t ld, char t, int64_t r0, int64_t c0) {
    return (t == 'N') ? x + r0 + c0 * ld : x + c0 + r0 * ld;
}

struct Operand {
    const double *host;   // caller pointer (host or device)
    int64_t ld;
    Kind kind;
    char t;               // 'N' or 'T' (how the stored matrix maps to op(X))
    int64_t srows, scols; // stored extents
};

// Uploads block (r0..r0+nr, c0..c0+nc) of op(X) (stored according to X.t) from
// the host into the device buffer dst (leading dimension dld, same stored
// orientation), at the block's own position if `at_origin` is false, else at
// the origin of dst.
int upload_op_block(cosma_amd_handle_t h, const Operand &X, double *dst, int64_t dld, int64_t r0,
                    int64_t c0, int64_t nr, int64_t nc, bool at_origin) {
    const int64_t sr0 = (X.t == 'N') ? r0 : c0, sc0 = (X.t == 'N') ? c0 : r0;
    const int64_t snr = (X.t == 'N') ? nr : nc, snc = (X.t == 'N') ? nc : nr;
    double *d = at_origin ? dst : dst + sr0 + sc0 * dld;
    return upload(h, d, dld, X.host + sr0 + sc0 * X.ld, X.ld, snr, snc, X.kind);
}

}  // namespace

void release_stream_state(cosma_amd_handle_t h) {
    hipSetDevice(h->device);
    for (DevBuf *b : {&h->dA, &h->dB, &h->dC, &h->dC0, &h->dX, &h->ozaki_ws, &h->ozaki_aux}) {
        if (b->ptr) hipFree(b->ptr);
        b->ptr = nullptr;
        b->bytes = 0;
    }
    if (h->ozaki_host.ptr) hipHostFree(h->ozaki_host.ptr);
    h->ozaki_host.ptr = nullptr;
    h->ozaki_host.bytes = 0;
    for (auto *v : {&h->h2d_slots, &h->d2h_slots}) {
        for (auto &s : *v) {
            if (s.ev) hipEventDestroy(s.ev);
            if (s.ptr) hipHostFree(s.ptr);
        }
        v->clear();
    }
    for (auto e : h->events) hipEventDestroy(e);
    h->events.clear();
    for (auto &e : h->ozaki_ev) {
        if (e) hipEventDestroy(e);
        e = nullptr;
    }
    h->events_used = 0;
    h->dev_c = nullptr;
}

// Visiting order of the nTm x nTn C tiles: squares growing from tile (0,0).
// After s block rows of op(A) and s block columns of op(B) have arrived, s*s
// tiles can run, which front-loads the work each upload enables.
static std::vector<std::pair<int64_t, int64_t>> shell_order(int64_t nTm, int64_t nTn) {
    std::vector<std::pair<int64_t, int64_t>> out;
    out.reserve((size_t)(nTm * nTn));
    for (int64_t s = 0; s < std::max(nTm, nTn); ++s) {
        if (s < nTm)
            for (int64_t j = 0; j < std::min(s, nTn); ++j) out.emplace_back(s, j);
        if (s < nTn)
            for (int64_t i = 0; i <= std::min(s, nTm - 1); ++i) out.emplace_back(i, s);
    }
    return out;
}

// Resident schedule: each block of A and B crosses the host link once, into a
// device copy of the whole operand, and is reused by every tile that needs it.
// The first floor(nTk/2) panels run panel-major over all tiles (the blocks of
// panel p enable every tile's product p, so few uploads start much compute);
// the remaining ceil(nTk/2) panels run tile by tile, each tile then taking its
// deferred beta * C and leaving for the host while later tiles compute, so
// the returns overlap the compute instead of trailing it.
static int stream_resident(cosma_amd_handle_t h, const Operand &A, const Operand &B, Kind kc, int64_t m, int64_t n,
                           int64_t k, double alpha, double beta, double *c, int64_t ldc, bool copy_c_back,
                           int64_t tm, int64_t tn, int64_t tk, int64_t nTm, int64_t nTn, int64_t nTk,
                           bool defer_c, int64_t R) {
    int rc;
    double *dA = const_cast<double *>(A.host), *dB = const_cast<double *>(B.host), *dC = c, *dC0 = nullptr;
    int64_t ldA = A.ld, ldB = B.ld, ldC = ldc;
    if (A.kind != DEVICE) {
        if ((rc = grow(h->dA, (size_t)A.srows * A.scols * 8))) return rc;
        dA = static_cast<double *>(h->dA.ptr);
        ldA = A.srows;
    }
    if (B.kind != DEVICE) {
        if ((rc = grow(h->dB, (size_t)B.srows * B.scols * 8))) return rc;
        dB = static_cast<double *>(h->dB.ptr);
        ldB = B.srows;
    }
    if (kc != DEVICE) {
        if ((rc = grow(h->dC, (size_t)m * n * 8))) return rc;
        dC = static_cast<double *>(h->dC.ptr);
        ldC = m;
    }
    if (defer_c) {
        if ((rc = grow(h->dC0, (size_t)R * tm * tn * 8))) return rc;
        dC0 = static_cast<double *>(h->dC0.ptr);
    }

    std::vector<hipEvent_t> evA((size_t)(nTm * nTk), nullptr), evB((size_t)(nTk * nTn), nullptr);
    std::vector<hipEvent_t> slot_free((size_t)R, nullptr);
    const auto order = shell_order(nTm, nTn);
    const int64_t tail = (nTk + 1) / 2, head = nTk - tail;
    // Products run on one stream in the order their blocks are uploaded (the
    // upload stream is FIFO, so a second compute stream would only split the
    // GPU between kernels whose inputs arrive in the same order).
    hipStream_t s = h->streams[0];
    const double beta0 = defer_c ? 0.0 : beta;   // beta of each tile's first product
    TRACE("resident: %lld panel-major panels, %lld per tile, C %s", (long long)head, (long long)tail,
          defer_c ? "deferred" : "in place");

    // The first product would wait for two whole blocks: upload them in
    // k-chunks and multiply chunk by chunk, so the compute starts after the
    // first chunks arrive.  Later products wait for the whole blocks.
    bool first_split = false;
    if (nTk > 0 && (A.kind != DEVICE || B.kind != DEVICE)) {
        const int64_t i = order[0].first, j = order[0].second;
        const int64_t mi = std::min(tm, m), nj = std::min(tn, n), kp = std::min(tk, k);
        const int64_t nsub = std::max<int64_t>(1, std::min<int64_t>(4, kp / 256));
        if (nsub > 1) {
            const int64_t sub = (kp + nsub - 1) / nsub;
            const int64_t r0 = i * tm, c0 = j * tn;
            hipEvent_t ev = nullptr;
            for (int64_t kk = 0; kk < kp; kk += sub) {
                const int64_t kq = std::min(sub, kp - kk);
                if (A.kind != DEVICE && (rc = upload_op_block(h, A, dA, ldA, r0, kk, mi, kq, false))) return rc;
                if (B.kind != DEVICE && (rc = upload_op_block(h, B, dB, ldB, kk, c0, kq, nj, false))) return rc;
                if ((rc = record(h, h->h2d, &ev))) return rc;
                CA_HIP(hipStreamWaitEvent(s, ev, 0));
                rc = launch_dgemm(A.t, B.t, mi, nj, kq, alpha, op_block(dA, ldA, A.t, r0, kk), ldA,
                                  op_block(dB, ldB, B.t, kk, c0), ldB, kk == 0 ? beta0 : 1.0,
                                  dC + r0 + c0 * ldC, ldC, s);
                if (rc) return rc;
            }
            if (A.kind != DEVICE) evA[(size_t)(i * nTk)] = ev;
            if (B.kind != DEVICE) evB[(size_t)j] = ev;
            first_split = true;
            TRACE("first product in %lld k-chunks of %lld", (long long)nsub, (long long)sub);
        }
    }

    // Product of panel p into tile (i, j); blocks are uploaded on their first
    // use and waited for through their upload events.
    auto product = [&](int64_t i, int64_t j, int64_t p) -> int {
        if (first_split && p == 0 && i == order[0].first && j == order[0].second) return COSMA_AMD_OK;
        const int64_t mi = std::min(tm, m - i * tm), nj = std::min(tn, n - j * tn);
        const int64_t r0 = i * tm, c0 = j * tn, k0 = p * tk, kp = std::min(tk, k - k0);
        int rc2;
        if (A.kind != DEVICE) {
            hipEvent_t &e = evA[(size_t)(i * nTk + p)];
            if (!e) {
                if ((rc2 = upload_op_block(h, A, dA, ldA, r0, k0, mi, kp, false))) return rc2;
                if ((rc2 = record(h, h->h2d, &e))) return rc2;
            }
            CA_HIP(hipStreamWaitEvent(s, e, 0));
        }
        if (B.kind != DEVICE) {
            hipEvent_t &e = evB[(size_t)(p * nTn + j)];
            if (!e) {
                if ((rc2 = upload_op_block(h, B, dB, ldB, k0, c0, kp, nj, false))) return rc2;
                if ((rc2 = record(h, h->h2d, &e))) return rc2;
            }
            CA_HIP(hipStreamWaitEvent(s, e, 0));
        }
        return launch_dgemm(A.t, B.t, mi, nj, kp, alpha, op_block(dA, ldA, A.t, r0, k0), ldA,
                            op_block(dB, ldB, B.t, k0, c0), ldB, p > 0 ? 1.0 : beta0, dC + r0 + c0 * ldC,
                            ldC, s);
    };

    for (int64_t p = 0; p < head; ++p) {
        for (const auto &ij : order) {
            if ((rc = product(ij.first, ij.second, p))) return rc;
        }
    }
    for (size_t t = 0; t < order.size(); ++t) {
        const int64_t i = order[t].first, j = order[t].second;
        const int64_t mi = std::min(tm, m - i * tm), nj = std::min(tn, n - j * tn);
        const int64_t r0 = i * tm, c0 = j * tn;
        double *Ct = dC + r0 + c0 * ldC;
        const double *Ch = c + r0 + c0 * ldc;
        TRACE("tile (%lld,%lld)", (long long)i, (long long)j);
        for (int64_t p = head; p < nTk; ++p) {
            if ((rc = product(i, j, p))) return rc;
        }
        if (nTk == 0) {   // k == 0 or alpha == 0 (C is on the host here)
            if (beta != 0.0) {
                if ((rc = upload(h, Ct, ldC, Ch, ldc, mi, nj, kc))) return rc;
                hipEvent_t ev;
                if ((rc = record(h, h->h2d, &ev))) return rc;
                CA_HIP(hipStreamWaitEvent(s, ev, 0));
            }
            if ((rc = launch_scale(mi, nj, beta, Ct, ldC, s))) return rc;
        }
        if (defer_c) {
            const size_t q = t % (size_t)R;
            double *slot = dC0 + q * (size_t)(tm * tn);
            if (slot_free[q]) CA_HIP(hipStreamWaitEvent(h->h2d, slot_free[q], 0));
            if ((rc = upload(h, slot, tm, Ch, ldc, mi, nj, kc))) return rc;
            hipEvent_t ev;
            if ((rc = record(h, h->h2d, &ev))) return rc;
            CA_HIP(hipStreamWaitEvent(s, ev, 0));
            if ((rc = launch_add_scaled(mi, nj, beta, slot, tm, Ct, ldC, s))) return rc;
            if ((rc = record(h, s, &slot_free[q]))) return rc;
        }
        if (kc != DEVICE && copy_c_back) {
            hipEvent_t done;
            if ((rc = record(h, s, &done))) return rc;
            CA_HIP(hipStreamWaitEvent(h->d2h, done, 0));
            if ((rc = download(h, const_cast<double *>(Ch), ldc, Ct, ldC, mi, nj, kc))) return rc;
        }
    }
    if (kc != DEVICE && !copy_c_back) {
        h->dev_c = dC;
        h->dev_c_ld = ldC;
    }
    return COSMA_AMD_OK;
}

// Bounded schedule (operands larger than the device budget): per tile, the
// panels of op(A) and op(B) pass through two device slots per operand whose
// reuse waits on the product that last read them; C tiles of host C use one
// slot per compute stream, reused once the tile has left for the host.
static int stream_bounded(cosma_amd_handle_t h, const Operand &A, const Operand &B, Kind kc, int64_t m, int64_t n,
                   int64_t k, double alpha, double beta, double *c, int64_t ldc,
                   int64_t tm, int64_t tn, int64_t tk, int64_t nTm, int64_t nTn, int64_t nTk) {
    const int S = (int)h->streams.size();
    int rc;
    const size_t slotA = (size_t)tm * tk * 8, slotB = (size_t)tk * tn * 8, slotC = (size_t)tm * tn * 8;
    if (A.kind != DEVICE && (rc = grow(h->dA, 2 * slotA))) return rc;
    if (B.kind != DEVICE && (rc = grow(h->dB, 2 * slotB))) return rc;
    if (kc != DEVICE && (rc = grow(h->dC, (size_t)S * slotC))) return rc;

    hipEvent_t slot_free_A[2] = {nullptr, nullptr}, slot_free_B[2] = {nullptr, nullptr};
    std::vector<hipEvent_t> c_slot_free(S, nullptr);
    int64_t ring = 0;
    for (int64_t i = 0; i < nTm; ++i) {
        for (int64_t jj = 0; jj < nTn; ++jj) {
            const int64_t j = (i % 2 == 0) ? jj : nTn - 1 - jj;     // snake order
            const int64_t t = i * nTn + jj;
            hipStream_t s = h->streams[t % S];
            const int64_t mi = std::min(tm, m - i * tm), nj = std::min(tn, n - j * tn);
            const int64_t r0 = i * tm, c0 = j * tn;
            double *Ct;
            int64_t ldCt;
            if (kc == DEVICE) {
                Ct = c + r0 + c0 * ldc;
                ldCt = ldc;
            } else {
                Ct = static_cast<double *>(h->dC.ptr) + (size_t)(t % S) * tm * tn;
                ldCt = tm;
                // the slot's previous tile must have left the device before it is reused
                if (c_slot_free[t % S]) {
                    CA_HIP(hipStreamWaitEvent(h->h2d, c_slot_free[t % S], 0));
                    CA_HIP(hipStreamWaitEvent(s, c_slot_free[t % S], 0));
                }
            }
            TRACE("tile (%lld,%lld) stream %d C at %p ld %lld", (long long)i, (long long)j, (int)(t % S),
                  (void *)Ct, (long long)ldCt);
            if (kc != DEVICE && beta != 0.0) {
                if ((rc = upload(h, Ct, ldCt, c + r0 + c0 * ldc, ldc, mi, nj, kc))) return rc;
                hipEvent_t ev;
                if ((rc = record(h, h->h2d, &ev))) return rc;
                CA_HIP(hipStreamWaitEvent(s, ev, 0));
            }
            if (nTk == 0 && (rc = launch_scale(mi, nj, beta, Ct, ldCt, s))) return rc;
            for (int64_t p = 0; p < nTk; ++p) {
                const int64_t k0 = p * tk, kp = std::min(tk, k - k0);
                const int q = ring % 2;
                const double *Ap, *Bp;
                int64_t ldAp = A.ld, ldBp = B.ld;
                if (A.kind == DEVICE) {
                    Ap = op_block(A.host, A.ld, A.t, r0, k0);
                } else {
                    double *slot = static_cast<double *>(h->dA.ptr) + (size_t)q * tm * tk;
                    ldAp = (A.t == 'N') ? mi : kp;
                    if (slot_free_A[q]) CA_HIP(hipStreamWaitEvent(h->h2d, slot_free_A[q], 0));
                    if ((rc = upload_op_block(h, A, slot, ldAp, r0, k0, mi, kp, true))) return rc;
                    hipEvent_t e;
                    if ((rc = record(h, h->h2d, &e))) return rc;
                    CA_HIP(hipStreamWaitEvent(s, e, 0));
                    Ap = slot;
                }
                if (B.kind == DEVICE) {
                    Bp = op_block(B.host, B.ld, B.t, k0, c0);
                } else {
                    double *slot = static_cast<double *>(h->dB.ptr) + (size_t)q * tk * tn;
                    ldBp = (B.t == 'N') ? kp : nj;
                    if (slot_free_B[q]) CA_HIP(hipStreamWaitEvent(h->h2d, slot_free_B[q], 0));
                    if ((rc = upload_op_block(h, B, slot, ldBp, k0, c0, kp, nj, true))) return rc;
                    hipEvent_t e;
                    if ((rc = record(h, h->h2d, &e))) return rc;
                    CA_HIP(hipStreamWaitEvent(s, e, 0));
                    Bp = slot;
                }
                rc = launch_dgemm(A.t, B.t, mi, nj, kp, alpha, Ap, ldAp, Bp, ldBp, p == 0 ? beta : 1.0,
                                  Ct, ldCt, s);
                if (rc) return rc;
                hipEvent_t e;
                if ((rc = record(h, s, &e))) return rc;
                if (A.kind != DEVICE) slot_free_A[q] = e;
                if (B.kind != DEVICE) slot_free_B[q] = e;
                ++ring;
            }
            if (kc != DEVICE) {
                hipEvent_t done;
                if ((rc = record(h, s, &done))) return rc;
                CA_HIP(hipStreamWaitEvent(h->d2h, done, 0));
                if ((rc = download(h, c + r0 + c0 * ldc, ldc, Ct, ldCt, mi, nj, kc))) return rc;
                hipEvent_t freed;
                if ((rc = record(h, h->d2h, &freed))) return rc;
                c_slot_free[t % S] = freed;
            }
        }
    }
    return COSMA_AMD_OK;
}

static int stream_dgemm_body(cosma_amd_handle_t h, char ta, char tb, int64_t m, int64_t n, int64_t k,
                             double alpha, const double *a, int64_t lda, const double *b, int64_t ldb,
                             double beta, double *c, int64_t ldc, bool copy_c_back) {
    if (m == 0 || n == 0) return COSMA_AMD_OK;
    const bool scale_only = (k == 0 || alpha == 0.0);
    Operand A{a, lda, scale_only ? DEVICE : classify(a), ta, ta == 'N' ? m : k, ta == 'N' ? k : m};
    Operand B{b, ldb, scale_only ? DEVICE : classify(b), tb, tb == 'N' ? k : n, tb == 'N' ? n : k};
    const Kind kc = classify(c);
    h->events_used = 0;
    h->dev_c = nullptr;

    if (A.kind == DEVICE && B.kind == DEVICE && kc == DEVICE) {
        h->last_mode = 0;
        int rc = launch_dgemm(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, h->streams[0]);
        if (rc) return rc;
        CA_HIP(hipStreamSynchronize(h->streams[0]));
        return COSMA_AMD_OK;
    }

    const int64_t tm = std::min(h->tile_m, m), tn = std::min(h->tile_n, n);
    const int64_t tk = std::max<int64_t>(1, std::min(h->tile_k, k));
    const int64_t nTm = (m + tm - 1) / tm, nTn = (n + tn - 1) / tn;
    const int64_t nTk = scale_only ? 0 : (k + tk - 1) / tk;
    // host C with beta != 0: beta * C is added after the tile's last panel
    // from a ring of R device slots, so that C uploads queue behind the
    // operand blocks that start the compute.
    const bool defer_c = kc != DEVICE && beta != 0.0 && nTk > 0;
    const int64_t R = std::min<int64_t>(4, nTm * nTn);

    size_t free_b = 0, total_b = 0;
    CA_HIP(hipMemGetInfo(&free_b, &total_b));
    const size_t limit = h->mem_limit > 0 ? (size_t)h->mem_limit : (size_t)(0.8 * (double)free_b);
    const size_t needA = A.kind == DEVICE ? 0 : (size_t)A.srows * A.scols * 8;
    const size_t needB = B.kind == DEVICE ? 0 : (size_t)B.srows * B.scols * 8;
    const size_t needC = kc == DEVICE ? 0 : (size_t)m * n * 8;
    const size_t needC0 = defer_c ? (size_t)R * tm * tn * 8 : 0;
    const bool resident = needA + needB + needC + needC0 <= limit;
    h->last_mode = resident ? 1 : 2;
    TRACE("dgemm_host %c%c m=%lld n=%lld k=%lld alpha=%g beta=%g kinds A=%d B=%d C=%d tiles %lldx%lldx%lld "
          "grid %lldx%lldx%lld mode=%s", ta, tb, (long long)m, (long long)n, (long long)k, alpha, beta,
          (int)A.kind, (int)B.kind, (int)kc, (long long)tm, (long long)tn, (long long)tk, (long long)nTm,
          (long long)nTn, (long long)nTk, resident ? "resident" : "bounded");
    if (!resident && kc != DEVICE && !copy_c_back) {
        set_error("host streaming: copy_c_back = false needs C to fit on the device "
                  "(%zu bytes over a budget of %zu)", needA + needB + needC, limit);
        return COSMA_AMD_ERR_UNSUPPORTED;
    }
    int rc = resident ? stream_resident(h, A, B, kc, m, n, k, alpha, beta, c, ldc, copy_c_back,
                                        tm, tn, tk, nTm, nTn, nTk, defer_c, R)
                      : stream_bounded(h, A, B, kc, m, n, k, alpha, beta, c, ldc,
                                       tm, tn, tk, nTm, nTn, nTk);
    if (rc) return rc;
    return sync_all(h);
}

int stream_dgemm(cosma_amd_handle_t h, char ta, char tb, int64_t m, int64_t n, int64_t k,
                 double alpha, const double *a, int64_t lda, const double *b, int64_t ldb,
                 double beta, double *c, int64_t ldc, bool copy_c_back) {
    for (auto &s : h->d2h_slots) s.out = nullptr;   // nothing of an earlier call is pending
    const int rc = stream_dgemm_body(h, ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, copy_c_back);
    if (rc) abandon(h);
    return rc;
}

// ------------------------------------------------------ grid-layout multiply
namespace {

// Checks that the blocks lie inside a rows x cols matrix, are well formed and
// tile it exactly: in bounds + total area == rows * cols + no two blocks
// overlap (sorted sweep by column start, then row start, per column band).
int check_tiling(const char *name, const cosma_amd_block *blk, int64_t nb, int64_t rows, int64_t cols) {
    if (nb < 0 || (nb > 0 && !blk)) {
        set_error("layout multiply: %s has an invalid block list", name);
        return COSMA_AMD_ERR_INVALID_ARG;
    }
    __int128 area = 0;
    for (int64_t i = 0; i < nb; ++i) {
        const cosma_amd_block &b = blk[i];
        if (b.rows < 0 || b.cols < 0 || b.row < 0 || b.col < 0 || b.row + b.rows > rows ||
            b.col + b.cols > cols) {
            set_error("layout multiply: block %lld of %s ([%lld, +%lld) x [%lld, +%lld)) is outside the "
                      "%lld x %lld matrix", (long long)i, name, (long long)b.row, (long long)b.rows,
                      (long long)b.col, (long long)b.cols, (long long)rows, (long long)cols);
            return COSMA_AMD_ERR_INVALID_ARG;
        }
        if (b.rows > 0 && b.cols > 0 && (!b.data || b.ld < b.rows)) {
            set_error("layout multiply: block %lld of %s has a null pointer or ld %lld < rows %lld",
                      (long long)i, name, (long long)b.ld, (long long)b.rows);
            return COSMA_AMD_ERR_INVALID_ARG;
        }
        area += (__int128)b.rows * b.cols;
    }
    if (area != (__int128)rows * cols) {
        set_error("layout multiply: the blocks of %s cover %lld elements, the %lld x %lld matrix has %lld",
                  name, (long long)area, (long long)rows, (long long)cols, (long long)(rows * cols));
        return COSMA_AMD_ERR_INVALID_ARG;
    }
    // Equal total area plus no overlap means an exact tiling.
    std::vector<int64_t> order;
    order.reserve(nb);
    for (int64_t i = 0; i < nb; ++i)
        if (blk[i].rows > 0 && blk[i].cols > 0) order.push_back(i);
    std::sort(order.begin(), order.end(), [&](int64_t x, int64_t y) {
        return blk[x].row != blk[y].row ? blk[x].row < blk[y].row : blk[x].col < blk[y].col;
    });
    for (size_t p = 0; p < order.size(); ++p) {
        const cosma_amd_block &x = blk[order[p]];
        for (size_t q = p + 1; q < order.size(); ++q) {
            const cosma_amd_block &y = blk[order[q]];
            if (y.row >= x.row + x.rows) break;   // sorted by row start: no later block overlaps x
            if (y.col < x.col + x.cols && x.col < y.col + y.cols) {
                set_error("layout multiply: blocks %lld and %lld of %s overlap", (long long)order[p],
                          (long long)order[q], name);
                return COSMA_AMD_ERR_INVALID_ARG;
            }
        }
    }
    return COSMA_AMD_OK;
}

// Gathers the blocks of one operand into the column-major device matrix dst
// (leading dimension ld): host blocks are uploaded on h2d, device blocks are
// copied by the block-copy kernel on stream s.
int gather_blocks(cosma_amd_handle_t h, const cosma_amd_block *blk, int64_t nb, double *dst, int64_t ld,
                  hipStream_t s) {
    for (int64_t i = 0; i < nb; ++i) {
        const cosma_amd_block &b = blk[i];
        if (b.rows == 0 || b.cols == 0) continue;
        const double *src = static_cast<const double *>(b.data);
        double *d = dst + b.row + b.col * ld;
        const Kind kind = classify(src);
        int rc = (kind == DEVICE) ? launch_block_transform('N', b.rows, b.cols, 1.0, src, b.ld, 0.0, d, ld, s)
                                  : upload(h, d, ld, src, b.ld, b.rows, b.cols, kind);
        if (rc) return rc;
    }
    return COSMA_AMD_OK;
}

}  // namespace

static int layout_dgemm_body(cosma_amd_handle_t h, char ta, char tb, int64_t m, int64_t n, int64_t k,
                             double alpha, const cosma_amd_block *A, int64_t na, const cosma_amd_block *B,
                             int64_t nb, double beta, const cosma_amd_block *C, int64_t nc) {
    const int64_t arows = ta == 'N' ? m : k, acols = ta == 'N' ? k : m;
    const int64_t brows = tb == 'N' ? k : n, bcols = tb == 'N' ? n : k;
    int rc;
    if ((rc = check_tiling("C", C, nc, m, n))) return rc;
    if (m == 0 || n == 0) return COSMA_AMD_OK;
    const bool scale_only = (k == 0 || alpha == 0.0);
    if (!scale_only) {
        if ((rc = check_tiling("A", A, na, arows, acols))) return rc;
        if ((rc = check_tiling("B", B, nb, brows, bcols))) return rc;
    }
    h->events_used = 0;
    h->dev_c = nullptr;
    h->last_mode = 3;
    hipStream_t s = h->streams[0];

    // ---- device copies: A, B as stored, X = op(A) op(B), staging for host C blocks
    bool host_c = false;
    for (int64_t i = 0; i < nc; ++i) host_c |= (C[i].rows > 0 && C[i].cols > 0 && classify(C[i].data) != DEVICE);
    const size_t needA = scale_only ? 0 : (size_t)arows * acols * 8, needB = scale_only ? 0 : (size_t)brows * bcols * 8;
    const size_t needX = (size_t)m * n * 8, needC = host_c ? needX : 0;
    size_t free_b = 0, total_b = 0;
    CA_HIP(hipMemGetInfo(&free_b, &total_b));
    const size_t limit = h->mem_limit > 0 ? (size_t)h->mem_limit : (size_t)(0.8 * (double)free_b);
    if (needA + needB + needX + needC > limit) {
        set_error("layout multiply: the gathered operands need %zu bytes of device memory, the budget is %zu",
                  needA + needB + needX + needC, limit);
        return COSMA_AMD_ERR_UNSUPPORTED;
    }
    if ((rc = grow(h->dX, needX))) return rc;
    double *dX = static_cast<double *>(h->dX.ptr);
    double *dC = nullptr;
    if (host_c) {
        if ((rc = grow(h->dC, needC))) return rc;
        dC = static_cast<double *>(h->dC.ptr);
    }
    TRACE("layout dgemm %c%c m=%lld n=%lld k=%lld blocks %lld/%lld/%lld host C %d", ta, tb, (long long)m,
          (long long)n, (long long)k, (long long)na, (long long)nb, (long long)nc, (int)host_c);

    // ---- C blocks that live on the host (read only if beta != 0) go up first
    if (host_c && beta != 0.0) {
        for (int64_t i = 0; i < nc; ++i) {
            const cosma_amd_block &b = C[i];
            if (b.rows == 0 || b.cols == 0) continue;
            const Kind kind = classify(b.data);
            if (kind == DEVICE) continue;
            rc = upload(h, dC + b.row + b.col * m, m, static_cast<const double *>(b.data), b.ld, b.rows, b.cols, kind);
            if (rc) return rc;
        }
    }

    // ---- X = op(A) op(B) with (alpha, beta) = (1, 0), or X = 0 for the scale-only case
    if (scale_only) {
        CA_HIP(hipMemsetAsync(dX, 0, needX, s));
        alpha = 0.0;
    } else {
        if ((rc = grow(h->dA, needA)) || (rc = grow(h->dB, needB))) return rc;
        double *dA = static_cast<double *>(h->dA.ptr), *dB = static_cast<double *>(h->dB.ptr);
        if ((rc = gather_blocks(h, A, na, dA, arows, s))) return rc;
        if ((rc = gather_blocks(h, B, nb, dB, brows, s))) return rc;
        hipEvent_t up;
        if ((rc = record(h, h->h2d, &up))) return rc;
        CA_HIP(hipStreamWaitEvent(s, up, 0));
        if ((rc = launch_dgemm(ta, tb, m, n, k, 1.0, dA, arows, dB, brows, 0.0, dX, m, s))) return rc;
    }
    {
        hipEvent_t up;
        if ((rc = record(h, h->h2d, &up))) return rc;   // host C blocks are on the device
        CA_HIP(hipStreamWaitEvent(s, up, 0));
    }

    // ---- epilogue per C block: C_blk <- alpha * X_blk + beta * C_blk
    for (int64_t i = 0; i < nc; ++i) {
        const cosma_amd_block &b = C[i];
        if (b.rows == 0 || b.cols == 0) continue;
        const double *x = dX + b.row + b.col * m;
        if (classify(b.data) == DEVICE) {
            rc = launch_block_transform('N', b.rows, b.cols, alpha, x, m, beta, static_cast<double *>(b.data),
                                        b.ld, s);
        } else {
            rc = launch_block_transform('N', b.rows, b.cols, alpha, x, m, beta, dC + b.row + b.col * m, m, s);
        }
        if (rc) return rc;
    }
    if (host_c) {
        hipEvent_t done;
        if ((rc = record(h, s, &done))) return rc;
        CA_HIP(hipStreamWaitEvent(h->d2h, done, 0));
        for (int64_t i = 0; i < nc; ++i) {
            const cosma_amd_block &b = C[i];
            if (b.rows == 0 || b.cols == 0) continue;
            const Kind kind = classify(b.data);
            if (kind == DEVICE) continue;
            rc = download(h, static_cast<double *>(b.data), b.ld, dC + b.row + b.col * m, m, b.rows, b.cols, kind);
            if (rc) return rc;
        }
    }
    return sync_all(h);
}

int layout_dgemm(cosma_amd_handle_t h, char ta, char tb, int64_t m, int64_t n, int64_t k, double alpha,
                 const cosma_amd_block *A, int64_t na, const cosma_amd_block *B, int64_t nb, double beta,
                 const cosma_amd_block *C, int64_t nc) {
    for (auto &s : h->d2h_slots) s.out = nullptr;   // nothing of an earlier call is pending
    const int rc = layout_dgemm_body(h, ta, tb, m, n, k, alpha, A, na, B, nb, beta, C, nc);
    if (rc) abandon(h);
    return rc;
}

}  // namespace cosma_amd
