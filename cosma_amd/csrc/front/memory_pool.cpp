// memory_pool (interface of src/cosma/memory_pool.hpp:10-70): a stack of
// buffers in one host allocation, page-locked when COSMA_GPU_MEMORY_PINNING
// is ON.  Growth follows src/cosma/memory_pool.cpp:116-150 (reserve times
// the amortization factor, buffers padded to COSMA_CPU_MEMORY_ALIGNMENT).
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <new>
#include <stdexcept>

#include <cosma/environment_variables.hpp>
#include <cosma/memory_pool.hpp>

#include "front.hpp"

namespace cosma {

template <typename T>
memory_pool<T>::memory_pool() {
    pinned_ = get_memory_pinning();
    amortization = get_memory_pool_amortization();
}

template <typename T>
memory_pool<T>::~memory_pool() {
    if (!data_) return;
    if (pinned_) cosma_amd_host_free(data_);
    else std::free(data_);
}

template <typename T>
size_t memory_pool<T>::padded(size_t n) const {
    static const int alignment = get_cosma_cpu_memory_alignment();
    if (alignment <= 0) return n;
    const size_t bytes = n * sizeof(T), a = (size_t)alignment;
    const size_t rem = bytes % a;
    return rem ? n + (a - rem) / sizeof(T) : n;
}

template <typename T>
void memory_pool<T>::grow_to(size_t capacity) {
    if (capacity <= capacity_) return;
    const size_t bytes = capacity * sizeof(T);
    void *p = nullptr;
    if (pinned_) {
        if (cosma_amd_host_alloc(bytes, &p) != COSMA_AMD_OK) p = nullptr;
    } else {
        const int alignment = get_cosma_cpu_memory_alignment();
        const size_t a = std::max<size_t>(64, alignment > 0 ? (size_t)alignment : 0);
        if (posix_memalign(&p, a, std::max<size_t>(bytes, 1)) != 0) p = nullptr;
    }
    if (!p) {
        std::cout << "COSMA (memory pool): not enough space for " << bytes
                  << " bytes. Try setting the CPU memory limit (see environment variable COSMA_CPU_MAX_MEMORY)."
                  << std::endl;
        throw std::bad_alloc();
    }
    if (data_) {
        // buffers taken since the last growth may reach past the old storage:
        // only what it holds is copied
        std::memcpy(p, data_, std::min(size_, capacity_) * sizeof(T));
        if (pinned_) cosma_amd_host_free(data_);
        else std::free(data_);
    }
    if (output_)
        std::cout << "COSMA (memory pool): " << (pinned_ ? "page-locked" : "pageable") << " storage grown from "
                  << capacity_ << " to " << capacity << " elements" << std::endl;
    data_ = static_cast<T *>(p);
    capacity_ = capacity;
    resized = true;
    already_pinned = false;
}

template <typename T>
size_t memory_pool<T>::get_buffer_id(size_t size) {
    if (size == 0) throw std::invalid_argument("memory_pool: buffer of size 0");
    const size_t offset = size_;
    size_ += padded(size);
    ++n_buffers_;
    return offset;
}

template <typename T>
T *memory_pool<T>::get_buffer_pointer(size_t id) {
    if (size_ > capacity_) grow_to(size_);
    if (id >= capacity_) throw std::out_of_range("memory_pool: buffer id beyond the pool");
    return data_ + id;
}

template <typename T>
void memory_pool<T>::free_buffer(T *ptr, size_t size) {
    const size_t n = padded(size);
    if (!data_ || ptr < data_ || ptr + n > data_ + size_)
        throw std::invalid_argument("memory_pool: free_buffer of memory not in the pool");
    const size_t offset = (size_t)(ptr - data_);
    --n_buffers_;
    if (offset + n != size_) {
        released_.emplace_back(offset, n);   // not on top: released when the buffers above go
        return;
    }
    size_ = offset;
    // buffers released earlier that are now on top
    for (bool again = true; again;) {
        again = false;
        for (size_t i = 0; i < released_.size(); ++i) {
            if (released_[i].first + released_[i].second == size_) {
                size_ = released_[i].first;
                released_.erase(released_.begin() + (long)i);
                again = true;
                break;
            }
        }
    }
}

template <typename T>
void memory_pool<T>::resize(size_t capacity) {
    grow_to(capacity);
    size_ = capacity;
    released_.clear();
}

template <typename T>
void memory_pool<T>::reset() {
    size_ = 0;
    n_buffers_ = 0;
    released_.clear();
    resized = false;
    already_pinned = false;
}

template <typename T>
T *memory_pool<T>::get_pool_pointer() {
    return data_;
}

template <typename T>
void memory_pool<T>::turn_on_output() {
    output_ = true;
}

template <typename T>
size_t memory_pool<T>::size() {
    return size_;
}

template <typename T>
void memory_pool<T>::reserve(std::vector<size_t> &buffer_sizes) {
    size_t total = 0;
    for (auto &b : buffer_sizes) {
        b = padded(b);
        total += b;
    }
    const size_t want = padded((size_t)std::ceil((double)total * amortization));
    if (want > 0 && size_ + want > capacity_) grow_to(size_ + want);
}

template <typename T>
void memory_pool<T>::pin(T *, std::size_t) {
    already_pinned = true;
}

template <typename T>
void memory_pool<T>::unpin_all() {
    already_pinned = false;
}

template class memory_pool<double>;

}  // namespace cosma
