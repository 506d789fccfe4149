/* XXH32 and XXH64 over one segment on a group of four lanes, from the
 * published xxHash specification: XXH32 keeps four 32-bit accumulators over
 * 16-byte stripes, XXH64 four 64-bit accumulators over 32-byte stripes,
 * then tail rounds and a final avalanche.
 *
 * A stripe's accumulator step depends on the step before it, so a segment
 * is a serial chain of stripes; what runs in parallel is the four
 * accumulators. Lane q of the group (lanes base .. base + 3) owns
 * accumulator q and reads word q of every stripe, so the four lanes read a
 * stripe as one contiguous access, kAhead stripes ahead. Every lane of the
 * group calls the function (it shuffles within the group); the hash is
 * valid on lane q == 0. A lane with nothing to hash passes len 0. */
#ifndef CUDEC_AMD_XXHASH_H
#define CUDEC_AMD_XXHASH_H

#include "common.h"

namespace cudec_amd {
namespace xxh {

constexpr uint32_t kP32_1 = 0x9E3779B1u, kP32_2 = 0x85EBCA77u, kP32_3 = 0xC2B2AE3Du,
                   kP32_4 = 0x27D4EB2Fu, kP32_5 = 0x165667B1u;
constexpr uint64_t kP64_1 = 0x9E3779B185EBCA87ull, kP64_2 = 0xC2B2AE3D27D4EB4Full,
                   kP64_3 = 0x165667B19E3779F9ull, kP64_4 = 0x85EBCA77C2B2AE63ull,
                   kP64_5 = 0x27D4EB2F165667C5ull;
constexpr uint32_t kGroup = 4;
constexpr uint32_t kAhead = 8; /* stripes loaded before they are folded in */

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
/* Lane base + q's value of v. */
__device__ __forceinline__ uint32_t group_lane(uint32_t v, uint32_t base, uint32_t q) {
    return (uint32_t)__shfl((int)v, (int)(base + q));
}
__device__ __forceinline__ uint32_t round32(uint32_t acc, uint32_t w) {
    return rotl32(acc + w * kP32_2, 13) * kP32_1;
}
__device__ __forceinline__ uint64_t round64(uint64_t acc, uint64_t w) {
    return rotl64(acc + w * kP64_2, 31) * kP64_1;
}
__device__ __forceinline__ uint64_t merge64(uint64_t h, uint64_t acc) {
    return (h ^ round64(0, acc)) * kP64_1 + kP64_4;
}

/* XXH32 of p[0, len) with `seed`; q = lane % 4, base = lane - q. */
__device__ __forceinline__ uint32_t xxh32_group(gcu8* p, uint64_t len, uint32_t seed, uint32_t q,
                                                uint32_t base) {
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
            acc = round32(acc, w[k]);
        }
    }
    for (; s < stripes; s++) {
        acc = round32(acc, load32(p + 16 * s + 4 * q));
    }
    const uint32_t a0 = group_lane(acc, base, 0), a1 = group_lane(acc, base, 1),
                   a2 = group_lane(acc, base, 2), a3 = group_lane(acc, base, 3);
    uint32_t h = 0;
    if (q == 0) {
        h = len >= 16 ? rotl32(a0, 1) + rotl32(a1, 7) + rotl32(a2, 12) + rotl32(a3, 18)
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
    }
    return h;
}

/* XXH64 of p[0, len) with `seed`; q = lane % 4, base = lane - q. */
__device__ __forceinline__ uint64_t xxh64_group(gcu8* p, uint64_t len, uint64_t seed, uint32_t q,
                                                uint32_t base) {
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
            acc = round64(acc, w[k]);
        }
    }
    for (; s < stripes; s++) {
        acc = round64(acc, load64(p + 32 * s + 8 * q));
    }
    const uint32_t lo = (uint32_t)acc, hi = (uint32_t)(acc >> 32);
    uint64_t a[4];
#pragma unroll
    for (uint32_t k = 0; k < 4; k++) {
        a[k] = ((uint64_t)group_lane(hi, base, k) << 32) | group_lane(lo, base, k);
    }
    uint64_t h = 0;
    if (q == 0) {
        if (len >= 32) {
            h = rotl64(a[0], 1) + rotl64(a[1], 7) + rotl64(a[2], 12) + rotl64(a[3], 18);
            h = merge64(h, a[0]);
            h = merge64(h, a[1]);
            h = merge64(h, a[2]);
            h = merge64(h, a[3]);
        } else {
            h = seed + kP64_5;
        }
        h += len;
        uint64_t t = 32 * stripes;
        for (; t + 8 <= len; t += 8) {
            h ^= round64(0, load64(p + t));
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
    }
    return h;
}

}  // namespace xxh
}  // namespace cudec_amd

#endif
