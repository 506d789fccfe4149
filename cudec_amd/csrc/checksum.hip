/* Device checksums for gfx950: XXH32 (the LZ4 frame format's header, block
 * and content checksums) and XXH64 (the Zstd frame format's content
 * checksum), many segments per launch, from the published xxHash
 * specification: XXH32 keeps four 32-bit accumulators over 16-byte stripes,
 * XXH64 four 64-bit accumulators over 32-byte stripes, then tail rounds and
 * a final avalanche.
 *
 * A stripe's accumulator step depends on the step before it, so one
 * segment is a serial chain of stripes; what parallelises is the four
 * accumulators of a stripe and the segments themselves. Four lanes take
 * one segment (lane q owns accumulator q and reads word q of every stripe,
 * so the four lanes read each stripe as one contiguous 16- or 32-byte
 * access, eight stripes ahead), and a wave takes sixteen segments. Lane 0
 * of the four gathers the accumulators and hashes the tail. */
#include "common.h"
#include "cudec_mi355x.h"

namespace cudec_amd {
namespace {

constexpr uint32_t kP32_1 = 0x9E3779B1u, kP32_2 = 0x85EBCA77u, kP32_3 = 0xC2B2AE3Du,
                   kP32_4 = 0x27D4EB2Fu, kP32_5 = 0x165667B1u;
constexpr uint64_t kP64_1 = 0x9E3779B185EBCA87ull, kP64_2 = 0xC2B2AE3D27D4EB4Full,
                   kP64_3 = 0x165667B19E3779F9ull, kP64_4 = 0x85EBCA77C2B2AE63ull,
                   kP64_5 = 0x27D4EB2F165667C5ull;
constexpr uint32_t kLanesPerSegment = 4;
constexpr uint32_t kAhead = 8; /* stripes loaded before they are folded in */
constexpr int kBlock = 256;

__device__ __forceinline__ uint32_t rotl32(uint32_t x, uint32_t r) {
    return (x << r) | (x >> (32 - r));
}
__device__ __forceinline__ uint64_t rotl64(uint64_t x, uint32_t r) {
    return (x << r) | (x >> (64 - r));
}

__device__ __forceinline__ uint32_t load32(gcu8* p) {
    uint32_t v;
    __builtin_memcpy(&v, (const void*)p, 4);
    return v;
}
__device__ __forceinline__ uint64_t load64(gcu8* p) {
    uint64_t v;
    __builtin_memcpy(&v, (const void*)p, 8);
    return v;
}

/* The other three lanes' values of v, within this lane's group of four
 * (a group is lanes 4g .. 4g + 3). */
__device__ __forceinline__ uint32_t quad_lane(uint32_t v, uint32_t base, uint32_t q) {
    return (uint32_t)__shfl((int)v, (int)(base + q));
}

__device__ __forceinline__ uint32_t xxh32_round(uint32_t acc, uint32_t w) {
    return rotl32(acc + w * kP32_2, 13) * kP32_1;
}
__device__ __forceinline__ uint64_t xxh64_round(uint64_t acc, uint64_t w) {
    return rotl64(acc + w * kP64_2, 31) * kP64_1;
}
__device__ __forceinline__ uint64_t xxh64_merge(uint64_t h, uint64_t acc) {
    return (h ^ xxh64_round(0, acc)) * kP64_1 + kP64_4;
}

/* XXH32 of segments i (sizes[i] bytes at ptrs[i]) into out[i]. */
__global__ void __launch_bounds__(kBlock)
    xxh32_batch(const void* const* ptrs, const size_t* sizes, size_t n, uint32_t seed,
                uint32_t* out) {
    const uint32_t lane = __lane_id();
    const uint32_t q = lane % kLanesPerSegment;
    const uint32_t base = lane - q;
    const size_t per_grid = (size_t)gridDim.x * (kBlock / kLanesPerSegment);
    for (size_t i = (size_t)blockIdx.x * (kBlock / kLanesPerSegment) + threadIdx.x / kLanesPerSegment;
         ; i += per_grid) {
        /* every lane of the wave runs the loop the same number of times
         * (the shuffles below need the whole group of four) */
        const bool live = i < n;
        if (__ballot(live) == 0) {
            break;
        }
        gcu8* p = live ? (gcu8*)ptrs[i] : nullptr;
        const uint64_t len = live ? sizes[i] : 0;
        const uint64_t stripes = len / 16;
        uint32_t acc = q == 0 ? seed + kP32_1 + kP32_2 : q == 1 ? seed + kP32_2 : q == 2 ? seed
                                                                                        : seed - kP32_1;
        uint64_t s = 0;
        for (; s + kAhead <= stripes; s += kAhead) {
            uint32_t w[kAhead];
#pragma unroll
            for (uint32_t k = 0; k < kAhead; k++) {
                w[k] = load32(p + 16 * (s + k) + 4 * q);
            }
#pragma unroll
            for (uint32_t k = 0; k < kAhead; k++) {
                acc = xxh32_round(acc, w[k]);
            }
        }
        for (; s < stripes; s++) {
            acc = xxh32_round(acc, load32(p + 16 * s + 4 * q));
        }
        const uint32_t a0 = quad_lane(acc, base, 0), a1 = quad_lane(acc, base, 1),
                       a2 = quad_lane(acc, base, 2), a3 = quad_lane(acc, base, 3);
        if (live && q == 0) {
            uint32_t h = len >= 16 ? rotl32(a0, 1) + rotl32(a1, 7) + rotl32(a2, 12) + rotl32(a3, 18)
                                   : seed + kP32_5;
            h += (uint32_t)len;
            uint64_t t = 16 * stripes;
            for (; t + 4 <= len; t += 4) {
                h = rotl32(h + load32(p + t) * kP32_3, 17) * kP32_4;
            }
            for (; t < len; t++) {
                h = rotl32(h + (uint32_t)p[t] * kP32_5, 11) * kP32_1;
            }
            h ^= h >> 15;
            h *= kP32_2;
            h ^= h >> 13;
            h *= kP32_3;
            h ^= h >> 16;
            out[i] = h;
        }
    }
}

/* XXH64 of segments i into out[i]. */
__global__ void __launch_bounds__(kBlock)
    xxh64_batch(const void* const* ptrs, const size_t* sizes, size_t n, uint64_t seed,
                uint64_t* out) {
    const uint32_t lane = __lane_id();
    const uint32_t q = lane % kLanesPerSegment;
    const uint32_t base = lane - q;
    const size_t per_grid = (size_t)gridDim.x * (kBlock / kLanesPerSegment);
    for (size_t i = (size_t)blockIdx.x * (kBlock / kLanesPerSegment) + threadIdx.x / kLanesPerSegment;
         ; i += per_grid) {
        const bool live = i < n;
        if (__ballot(live) == 0) {
            break;
        }
        gcu8* p = live ? (gcu8*)ptrs[i] : nullptr;
        const uint64_t len = live ? sizes[i] : 0;
        const uint64_t stripes = len / 32;
        uint64_t acc = q == 0 ? seed + kP64_1 + kP64_2 : q == 1 ? seed + kP64_2 : q == 2 ? seed
                                                                                        : seed - kP64_1;
        uint64_t s = 0;
        for (; s + kAhead <= stripes; s += kAhead) {
            uint64_t w[kAhead];
#pragma unroll
            for (uint32_t k = 0; k < kAhead; k++) {
                w[k] = load64(p + 32 * (s + k) + 8 * q);
            }
#pragma unroll
            for (uint32_t k = 0; k < kAhead; k++) {
                acc = xxh64_round(acc, w[k]);
            }
        }
        for (; s < stripes; s++) {
            acc = xxh64_round(acc, load64(p + 32 * s + 8 * q));
        }
        const uint32_t lo = (uint32_t)acc, hi = (uint32_t)(acc >> 32);
        uint64_t a[4];
#pragma unroll
        for (uint32_t k = 0; k < 4; k++) {
            a[k] = ((uint64_t)quad_lane(hi, base, k) << 32) | quad_lane(lo, base, k);
        }
        if (live && q == 0) {
            uint64_t h;
            if (len >= 32) {
                h = rotl64(a[0], 1) + rotl64(a[1], 7) + rotl64(a[2], 12) + rotl64(a[3], 18);
                h = xxh64_merge(h, a[0]);
                h = xxh64_merge(h, a[1]);
                h = xxh64_merge(h, a[2]);
                h = xxh64_merge(h, a[3]);
            } else {
                h = seed + kP64_5;
            }
            h += len;
            uint64_t t = 32 * stripes;
            for (; t + 8 <= len; t += 8) {
                h ^= xxh64_round(0, load64(p + t));
                h = rotl64(h, 27) * kP64_1 + kP64_4;
            }
            if (t + 4 <= len) {
                h ^= (uint64_t)load32(p + t) * kP64_1;
                h = rotl64(h, 23) * kP64_2 + kP64_3;
                t += 4;
            }
            for (; t < len; t++) {
                h ^= (uint64_t)p[t] * kP64_5;
                h = rotl64(h, 11) * kP64_1;
            }
            h ^= h >> 33;
            h *= kP64_2;
            h ^= h >> 29;
            h *= kP64_3;
            h ^= h >> 32;
            out[i] = h;
        }
    }
}

cudec_status checked_launch() {
    return hipGetLastError() == hipSuccess ? CUDEC_OK : CUDEC_ERR_CUDA;
}

unsigned grid_for(size_t n) {
    const size_t per_block = kBlock / kLanesPerSegment;
    size_t g = (n + per_block - 1) / per_block;
    return (unsigned)(g > 8192 ? 8192 : g);
}

}  // namespace
}  // namespace cudec_amd

extern "C" cudec_status cudec_mi355x_xxh32_batch(const void* const* d_ptrs, const size_t* d_sizes,
                                                 size_t count, uint32_t seed, uint32_t* d_hashes,
                                                 cudec_stream_t stream) {
    using namespace cudec_amd;
    if (d_ptrs == nullptr || d_sizes == nullptr || d_hashes == nullptr || count == 0) {
        return CUDEC_ERR_INVALID_ARGUMENT;
    }
    (void)hipGetLastError();
    hipLaunchKernelGGL(xxh32_batch, dim3(grid_for(count)), dim3(kBlock), 0,
                       reinterpret_cast<hipStream_t>(stream), d_ptrs, d_sizes, count, seed,
                       d_hashes);
    return checked_launch();
}

extern "C" cudec_status cudec_mi355x_xxh64_batch(const void* const* d_ptrs, const size_t* d_sizes,
                                                 size_t count, uint64_t seed, uint64_t* d_hashes,
                                                 cudec_stream_t stream) {
    using namespace cudec_amd;
    if (d_ptrs == nullptr || d_sizes == nullptr || d_hashes == nullptr || count == 0) {
        return CUDEC_ERR_INVALID_ARGUMENT;
    }
    (void)hipGetLastError();
    hipLaunchKernelGGL(xxh64_batch, dim3(grid_for(count)), dim3(kBlock), 0,
                       reinterpret_cast<hipStream_t>(stream), d_ptrs, d_sizes, count, seed,
                       d_hashes);
    return checked_launch();
}
