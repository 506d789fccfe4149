/* Ownership of HIP runtime objects for the host-side paths (streaming
 * context, frame decoder, serving loop, tuner). Internal.
 *
 * Handle<H, Release> holds one runtime handle (a stream, an event, a
 * device or pinned allocation) and hands it back through Release when it
 * is replaced or goes out of scope; it moves but does not copy.
 * Region<Space> is a grow-only byte area in one memory space, built on a
 * Handle: reserve(n) keeps the area while it already holds n bytes and
 * otherwise replaces it with one of exactly n bytes. */
#ifndef CUDEC_AMD_HIP_OWNER_H
#define CUDEC_AMD_HIP_OWNER_H

#include <hip/hip_runtime.h>

#include <cstddef>
#include <utility>

namespace cudec_amd {

template <class H, hipError_t (*Release)(H)>
class Handle {
public:
    Handle() = default;
    Handle(const Handle&) = delete;
    Handle& operator=(const Handle&) = delete;
    Handle(Handle&& other) noexcept : h_(std::exchange(other.h_, H{})) {}
    Handle& operator=(Handle&& other) noexcept {
        if (this != &other) {
            adopt(std::exchange(other.h_, H{}));
        }
        return *this;
    }
    ~Handle() { adopt(H{}); }

    H get() const { return h_; }
    explicit operator bool() const { return h_ != H{}; }

    /* Gives back the current handle and takes h. */
    void adopt(H h) {
        if (h_ != H{}) {
            (void)Release(h_);
        }
        h_ = h;
    }

    /* Fills the handle from a runtime call that writes one, such as
     * hipEventCreate(&h): the old handle is given back first, and nothing
     * is kept when the call fails. */
    template <class Make>
    hipError_t make(Make&& call) {
        adopt(H{});
        H fresh{};
        const hipError_t e = call(&fresh);
        if (e == hipSuccess) {
            h_ = fresh;
        }
        return e;
    }

private:
    H h_{};
};

using Stream = Handle<hipStream_t, hipStreamDestroy>;
using Event = Handle<hipEvent_t, hipEventDestroy>;

/* A stream that does not synchronise with the null stream, at the default
 * priority. */
inline hipError_t make_stream(Stream& s) {
    return s.make([](hipStream_t* h) { return hipStreamCreateWithFlags(h, hipStreamNonBlocking); });
}

/* An event used only for ordering and completion (no timing). */
inline hipError_t make_sync_event(Event& e) {
    return e.make([](hipEvent_t* h) { return hipEventCreateWithFlags(h, hipEventDisableTiming); });
}

/* An event that records timestamps. */
inline hipError_t make_timing_event(Event& e) {
    return e.make([](hipEvent_t* h) { return hipEventCreate(h); });
}

/* Memory spaces a Region can live in: how to get n bytes and give them back. */
struct DeviceSpace {
    static hipError_t get(void** p, size_t n) { return hipMalloc(p, n); }
    static hipError_t put(void* p) { return hipFree(p); }
};
struct PinnedSpace {
    static hipError_t get(void** p, size_t n) { return hipHostMalloc(p, n, hipHostMallocDefault); }
    static hipError_t put(void* p) { return hipHostFree(p); }
};

template <class Space>
class Region {
public:
    void* data() const { return mem_.get(); }
    size_t capacity() const { return bytes_; }

    template <class T>
    T* as() const {
        return static_cast<T*>(mem_.get());
    }

    hipError_t reserve(size_t n) {
        if (mem_ && n <= bytes_) {
            return hipSuccess;
        }
        bytes_ = 0;
        /* a zero-byte request still gets a distinct, valid address */
        const hipError_t e = mem_.make([n](void** p) { return Space::get(p, n == 0 ? 1 : n); });
        if (e == hipSuccess) {
            bytes_ = n;
        }
        return e;
    }

private:
    Handle<void*, Space::put> mem_;
    size_t bytes_ = 0;
};

using DeviceRegion = Region<DeviceSpace>;
using PinnedRegion = Region<PinnedSpace>;

}  // namespace cudec_amd

#endif /* CUDEC_AMD_HIP_OWNER_H */
