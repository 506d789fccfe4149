/* Shared definitions for the gfx950 batch decoders. Internal. */
#ifndef CUDEC_AMD_COMMON_H
#define CUDEC_AMD_COMMON_H

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "cudec.h"

namespace cudec_amd {

constexpr unsigned kWave = 64;

/* One chunk per wave and at most INT32_MAX waves in a grid: the largest
 * chunk_count any entry point maps onto a launch (the wave64 restatement of
 * src/batch_limits.h:24-28). Above it a call is refused, never truncated. */
constexpr size_t kMaxBatchChunks = static_cast<size_t>(INT32_MAX) * kWave;
static_assert(kMaxBatchChunks < SIZE_MAX, "SIZE_MAX must stay over the limit");

enum Format : int { kFormatLz4 = 0, kFormatSnappy = 1 };

__device__ __forceinline__ uint32_t uniform32(uint32_t v) {
    return __builtin_amdgcn_readfirstlane(v);
}

__device__ __forceinline__ uint64_t uniform64(uint64_t v) {
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return ((uint64_t)hi << 32) | lo;
}

template <class T>
__device__ __forceinline__ T* uniform_ptr(T* p) {
    return reinterpret_cast<T*>(uniform64(reinterpret_cast<uintptr_t>(p)));
}

/* Global-memory byte pointers. Chunk pointers arrive through tables, so the
 * compiler sees generic pointers and would emit flat instructions, which
 * count against the LDS wait counter too: every wait for an LDS read would
 * also wait for outstanding global loads and stores. Casting them into the
 * global address space (a no-op on the address) gives global_* instead. */
typedef __attribute__((address_space(1))) uint8_t gu8;
typedef __attribute__((address_space(1))) const uint8_t gcu8;
typedef __attribute__((address_space(1))) const uint32_t gcu32;

/* A chunk pointer from a table, as a wave-uniform global pointer. */
__device__ __forceinline__ gu8* global_bytes(void* p) {
    return (gu8*)uniform64(reinterpret_cast<uintptr_t>(p));
}
__device__ __forceinline__ gcu8* global_bytes(const void* p) {
    return (gcu8*)uniform64(reinterpret_cast<uintptr_t>(p));
}

/* 16 bytes from / to global memory at any alignment. */
__device__ __forceinline__ uint4 gload16(gcu8* p) {
    uint4 v;
    __builtin_memcpy(&v, (const void*)p, 16);
    return v;
}
__device__ __forceinline__ void gstore16(gu8* p, const uint4& v) {
    __builtin_memcpy((void*)p, &v, 16);
}

/* Orders this wave's global-memory writes before its later reads of the
 * same bytes, across lanes. Within one wavefront the hardware issues and
 * completes vector memory operations in program order through one L1, so
 * a wavefront-scope fence needs no wait; it stops the compiler from moving
 * a load above an earlier store (the role __syncwarp plays in
 * src/lz4_decode.cuh:109-110,143-144). */
__device__ __forceinline__ void wave_order() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ void write_result(cudec_chunk_result* results,
                                             uint64_t chunk, int status,
                                             uint64_t bytes) {
    cudec_chunk_result r;
    r.status = status;
    r.reserved = 0;
    r.bytes_written = status == CUDEC_OK ? bytes : 0;
    results[chunk] = r;
}

/* Host-side launch entry points of each kernel form (defined in the .hip
 * files). Each returns the hipError_t of its launch. */
hipError_t launch_plain(int format, const void* const* srcs,
                        const size_t* src_sizes, void* const* dsts,
                        const size_t* caps, size_t n,
                        cudec_chunk_result* results, hipStream_t stream);

/* The current device's decode plan of one format (tuned.hip): for each of
 * the 256 cost bins (cost = capacity + (capacity - compressed size), four
 * bins per octave, bin 0 the costliest), the lanes that decode a chunk of
 * that bin, 128 (two pipelined waves) or 64 (one wave). A threshold sets
 * the plan that puts chunks costing at least min_cost on two waves (0:
 * every chunk, UINT64_MAX none); get_pipe_min_cost reports the threshold
 * last set. */
constexpr int kCostBins = 256;
hipError_t set_pipe_min_cost(int format, uint64_t min_cost);
hipError_t get_pipe_min_cost(int format, uint64_t* min_cost);
hipError_t set_decode_plan(int format, const uint16_t* lanes);
hipError_t get_decode_plan(int format, uint16_t* lanes);
uint32_t host_cost_bin(uint64_t src_size, uint64_t cap);

}  // namespace cudec_amd

#endif /* CUDEC_AMD_COMMON_H */
