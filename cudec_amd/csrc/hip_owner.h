/* Small owners for HIP runtime resources used by the host-side paths
 * (streaming context, frame decoder, serving loop). Internal. */
#ifndef CUDEC_AMD_HIP_OWNER_H
#define CUDEC_AMD_HIP_OWNER_H

#include <hip/hip_runtime.h>

#include <cstddef>

namespace cudec_amd {

/* Device memory, grow-only: ensure(n) keeps the allocation when it already
 * holds n bytes, else frees and allocates n bytes. */
struct DeviceBuffer {
    void* p = nullptr;
    size_t size = 0;
    DeviceBuffer() = default;
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    ~DeviceBuffer() { release(); }
    void release() {
        if (p != nullptr) {
            (void)hipFree(p);
        }
        p = nullptr;
        size = 0;
    }
    hipError_t ensure(size_t n) {
        if (n <= size && p != nullptr) {
            return hipSuccess;
        }
        release();
        const hipError_t e = hipMalloc(&p, n == 0 ? 1 : n);
        if (e != hipSuccess) {
            p = nullptr;
            return e;
        }
        size = n;
        return hipSuccess;
    }
};

/* Pinned host memory (hipHostMalloc), grow-only like DeviceBuffer. */
struct PinnedBuffer {
    void* p = nullptr;
    size_t size = 0;
    PinnedBuffer() = default;
    PinnedBuffer(const PinnedBuffer&) = delete;
    PinnedBuffer& operator=(const PinnedBuffer&) = delete;
    ~PinnedBuffer() { release(); }
    void release() {
        if (p != nullptr) {
            (void)hipHostFree(p);
        }
        p = nullptr;
        size = 0;
    }
    hipError_t ensure(size_t n) {
        if (n <= size && p != nullptr) {
            return hipSuccess;
        }
        release();
        const hipError_t e = hipHostMalloc(&p, n == 0 ? 1 : n, hipHostMallocDefault);
        if (e != hipSuccess) {
            p = nullptr;
            return e;
        }
        size = n;
        return hipSuccess;
    }
};

/* A non-blocking stream at the default priority. */
struct StreamOwner {
    hipStream_t s = nullptr;
    StreamOwner() = default;
    StreamOwner(const StreamOwner&) = delete;
    StreamOwner& operator=(const StreamOwner&) = delete;
    ~StreamOwner() {
        if (s != nullptr) {
            (void)hipStreamDestroy(s);
        }
    }
    hipError_t create() { return hipStreamCreateWithFlags(&s, hipStreamNonBlocking); }
};

struct EventOwner {
    hipEvent_t e = nullptr;
    EventOwner() = default;
    EventOwner(const EventOwner&) = delete;
    EventOwner& operator=(const EventOwner&) = delete;
    ~EventOwner() {
        if (e != nullptr) {
            (void)hipEventDestroy(e);
        }
    }
    hipError_t create() { return hipEventCreateWithFlags(&e, hipEventDisableTiming); }
};

}  // namespace cudec_amd

#endif /* CUDEC_AMD_HIP_OWNER_H */
