// hs::Queue stands in for the reference's sycl::queue (main.cpp:189-192):
// a GPU queue owns one libhs_amd handle (device + ordered HIP stream); a CPU
// queue is the host side and launches nothing. PinnedAllocator stands in for
// sycl::usm_allocator<T, usm::alloc::host> (SymmetricMatrix.hpp:53).
#pragma once

#include <cstddef>
#include <new>
#include <stdexcept>
#include <string>

#include "hs_abi.h"

namespace hs {

struct cpu_selector_t {};
struct gpu_selector_t {};
inline constexpr cpu_selector_t cpu_selector_v{};
inline constexpr gpu_selector_t gpu_selector_v{};

// Throws std::runtime_error with the library's message for a failed call.
void check(int status, const char* what);

class Queue {
public:
    explicit Queue(gpu_selector_t, int device = 0);
    explicit Queue(cpu_selector_t);
    ~Queue();
    Queue(const Queue&) = delete;
    Queue& operator=(const Queue&) = delete;

    bool is_gpu() const { return handle_ != nullptr; }
    hs_handle_t handle() const;
    void wait();
    // Sizes the device workspace for a system of order N in blocks of b (M
    // test points of a GP prediction): the set-up step before any solver
    // call on this queue, which never allocates by itself.
    void reserve(std::size_t N, int b, std::size_t M = 0);
    std::size_t global_mem_size() const;
    std::string name() const;

private:
    hs_handle_t handle_ = nullptr;
    int device_ = -1;
};

// True when the HIP runtime sees a device (decided once per process).
bool pinned_host_available();

// Pinned host memory when a GPU runtime is present, ordinary aligned memory
// otherwise (host-only tools such as the text parser still work).
template <typename T>
struct PinnedAllocator {
    typedef T value_type;
    PinnedAllocator() = default;
    template <typename U>
    PinnedAllocator(const PinnedAllocator<U>&) {}
    T* allocate(std::size_t n) {
        if (pinned_host_available()) {
            void* p = nullptr;
            if (hs_host_malloc(&p, n * sizeof(T)) != HS_OK || !p) throw std::bad_alloc();
            return static_cast<T*>(p);
        }
        return static_cast<T*>(::operator new(n * sizeof(T), std::align_val_t(64)));
    }
    void deallocate(T* p, std::size_t) {
        if (pinned_host_available()) hs_host_free(p);
        else ::operator delete(p, std::align_val_t(64));
    }
};
template <typename T, typename U>
bool operator==(const PinnedAllocator<T>&, const PinnedAllocator<U>&) { return true; }
template <typename T, typename U>
bool operator!=(const PinnedAllocator<T>&, const PinnedAllocator<U>&) { return false; }

// Device buffer owned by one queue.
template <typename T>
class DeviceArray {
public:
    DeviceArray() = default;
    explicit DeviceArray(std::size_t n) { resize(n); }
    ~DeviceArray() { release(); }
    DeviceArray(const DeviceArray&) = delete;
    DeviceArray& operator=(const DeviceArray&) = delete;
    void resize(std::size_t n) {
        release();
        void* p = nullptr;
        check(hs_malloc(&p, n * sizeof(T)), "hs_malloc");
        ptr_ = static_cast<T*>(p);
        n_ = n;
    }
    void release() {
        if (ptr_) hs_free(ptr_);
        ptr_ = nullptr;
        n_ = 0;
    }
    T* data() const { return ptr_; }
    std::size_t size() const { return n_; }

private:
    T* ptr_ = nullptr;
    std::size_t n_ = 0;
};

}  // namespace hs
