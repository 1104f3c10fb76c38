// SyncedMemory.h — host/device buffer pair with lazy synchronisation.
//
// Same API as src/SyncedMemory.h of johnjohnlin/UMI:
//   SyncedMemory<T>: non-owning (cpu, gpu, n) view; get_{cpu,gpu}_wo mark a
//     side as newest, get_{cpu,gpu}_ro copy from the other side first when it
//     is newer, get_*_rw = ro then wo               (src/SyncedMemory.h:5-56)
//   MemoryBuffer<T>: owns n elements on the host and on the device;
//     CreateSync(n, offset) returns a view          (src/SyncedMemory.h:58-116)
// Device memory and copies go through the libumi_hip C ABI.  Two defects of
// the reference are not reproduced: MemoryBuffer::Free releases its memory
// (src/SyncedMemory.h:72 tests the wrong pointer) and the move operations
// carry the element count and state (src/SyncedMemory.h:33-48).
#pragma once
#include <cstddef>
#include <cstdio>
#include <cstdlib>

#include "umi_hip.h"

namespace umi_detail {
inline void check(umi_status st, const char *what) {
    if (st != UMI_OK) {
        std::fprintf(stderr, "%s failed (status %d): %s\n", what, st, umi_last_error());
        std::abort();
    }
}
}  // namespace umi_detail

template <typename T>
class SyncedMemory {
    enum MemoryState { GPU_IS_NEW, CPU_IS_NEW, SYNCED } state_ = SYNCED;
    T *cpu_ = nullptr, *gpu_ = nullptr;
    size_t n_ = 0;

    void Sync() {
        if (state_ == CPU_IS_NEW)
            umi_detail::check(umi_memcpy(gpu_, cpu_, sizeof(T) * n_, UMI_H2D), "SyncedMemory host->device copy");
        else if (state_ == GPU_IS_NEW)
            umi_detail::check(umi_memcpy(cpu_, gpu_, sizeof(T) * n_, UMI_D2H), "SyncedMemory device->host copy");
        state_ = SYNCED;
    }

   public:
    void Reset(T *cpu, T *gpu, const size_t n) {
        cpu_ = cpu;
        gpu_ = gpu;
        n_ = n;
        state_ = SYNCED;
    }
    SyncedMemory() = default;
    SyncedMemory(T *cpu, T *gpu, const size_t n) { Reset(cpu, gpu, n); }
    SyncedMemory &operator=(const SyncedMemory &) = delete;
    SyncedMemory(const SyncedMemory &) = delete;
    SyncedMemory &operator=(SyncedMemory &&o) {
        if (this != &o) {
            cpu_ = o.cpu_;
            gpu_ = o.gpu_;
            n_ = o.n_;
            state_ = o.state_;
            o.cpu_ = o.gpu_ = nullptr;
            o.n_ = 0;
        }
        return *this;
    }
    SyncedMemory(SyncedMemory &&o) { *this = static_cast<SyncedMemory &&>(o); }

    T *get_cpu_wo() {
        state_ = CPU_IS_NEW;
        return cpu_;
    }
    T *get_gpu_wo() {
        state_ = GPU_IS_NEW;
        return gpu_;
    }
    const T *get_cpu_ro() {
        if (state_ == GPU_IS_NEW) Sync();
        return cpu_;
    }
    const T *get_gpu_ro() {
        if (state_ == CPU_IS_NEW) Sync();
        return gpu_;
    }
    T *get_cpu_rw() {
        get_cpu_ro();
        return get_cpu_wo();
    }
    T *get_gpu_rw() {
        get_gpu_ro();
        return get_gpu_wo();
    }
    size_t size() const { return n_; }
};

template <typename T>
class MemoryBuffer {
    T *cpu_ = nullptr, *gpu_ = nullptr;
    void Alloc(const size_t n) {
        if (n == 0) return;
        cpu_ = new T[n];
        void *p = nullptr;
        umi_detail::check(umi_malloc(&p, sizeof(T) * n), "MemoryBuffer device allocation");
        gpu_ = static_cast<T *>(p);
    }

   public:
    void Free() {
        delete[] cpu_;
        if (gpu_) umi_free(gpu_);
        cpu_ = gpu_ = nullptr;
    }
    void Realloc(const size_t n) {
        Free();
        Alloc(n);
    }
    MemoryBuffer(const size_t n = 0) { Alloc(n); }
    ~MemoryBuffer() { Free(); }
    MemoryBuffer &operator=(const MemoryBuffer &) = delete;
    MemoryBuffer(const MemoryBuffer &) = delete;
    MemoryBuffer &operator=(MemoryBuffer &&o) {
        if (this != &o) {
            Free();
            cpu_ = o.cpu_;
            gpu_ = o.gpu_;
            o.cpu_ = o.gpu_ = nullptr;
        }
        return *this;
    }
    MemoryBuffer(MemoryBuffer &&o) : cpu_(o.cpu_), gpu_(o.gpu_) { o.cpu_ = o.gpu_ = nullptr; }

    SyncedMemory<T> CreateSync(const size_t n, const size_t offset = 0) {
        return SyncedMemory<T>(cpu_ + offset, gpu_ + offset, n);
    }
};
