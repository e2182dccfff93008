// memory_pool: the host memory of the COSMA interface's matrices.
//
// Interface of src/cosma/memory_pool.hpp:10-70 (implementation
// src/cosma/memory_pool.cpp): one contiguous pool used as a stack of
// buffers.  A buffer is identified by its offset in the pool
// (get_buffer_id); its address is asked for when it is used
// (get_buffer_pointer), because growing the pool moves it.  reserve() sizes
// the pool for a set of buffers times COSMA_MEMORY_POOL_AMORTIZATION, so
// that later growth is rare; every buffer is padded to a multiple of
// COSMA_CPU_MEMORY_ALIGNMENT bytes when that is set (> 0), and the pool
// itself is aligned to it.
//
// Unlike the reference, whose pool is a std::vector registered with the GPU
// runtime range by range (pin(), src/cosma/pinned_buffers.cpp:6-47), the
// storage here is allocated page-locked (hipHostMalloc) when
// COSMA_GPU_MEMORY_PINNING is ON (the default): the host-streaming GEMM then
// moves every buffer by DMA, and pin() / unpin_all() have nothing to
// register.  With the knob OFF the storage is ordinary pageable memory,
// which the GEMM stages through its own pinned buffers.
#pragma once
#include <cstddef>
#include <vector>

namespace cosma {

template <typename T>
class memory_pool {
  public:
    memory_pool();
    ~memory_pool();
    memory_pool(const memory_pool &) = delete;
    memory_pool &operator=(const memory_pool &) = delete;

    // Offset of a new buffer of `size` elements on top of the stack.
    size_t get_buffer_id(size_t size);
    // Address of buffer `id`; grows the storage (keeping its contents) if
    // buffers were added beyond the current capacity.
    T *get_buffer_pointer(size_t id);
    // Releases a buffer.  Buffers are released from the top of the stack; one
    // released out of order is kept until the buffers above it are released.
    void free_buffer(T *ptr, size_t size);

    // Storage of exactly `capacity` elements, all of them in use.
    void resize(size_t capacity);
    // Forgets every buffer (the storage is kept for reuse).
    void reset();

    T *get_pool_pointer();
    void turn_on_output();
    // Elements in use (top of the stack).
    size_t size();
    // Makes room for these buffers (each padded to the alignment, in place)
    // times the amortization factor.
    void reserve(std::vector<size_t> &buffer_sizes);

    // The storage is page-locked already (COSMA_GPU_MEMORY_PINNING=ON) or is
    // staged by the GEMM (OFF): nothing to register.
    void pin(T *ptr, std::size_t size);
    void unpin_all();

    bool already_pinned = false;
    bool resized = false;
    // scaling factor for the buffer growth (COSMA_MEMORY_POOL_AMORTIZATION)
    double amortization = 1.2;

    // Not in the reference interface: capacity in elements, whether the
    // storage is page-locked, and the alignment padding of a size.
    size_t capacity() const { return capacity_; }
    bool page_locked() const { return pinned_; }
    size_t padded(size_t n) const;

  private:
    void grow_to(size_t capacity);
    T *data_ = nullptr;
    size_t size_ = 0, capacity_ = 0, n_buffers_ = 0;
    bool pinned_ = true;
    bool output_ = false;
    // buffers released out of order: (offset, padded size)
    std::vector<std::pair<size_t, size_t>> released_;
};

}  // namespace cosma
