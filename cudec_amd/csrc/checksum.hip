/* Device checksums for gfx950: XXH32 (the LZ4 frame format's header, block
 * and content checksums) and XXH64 (the Zstd frame format's content
 * checksum), many segments per launch. Four lanes take one segment (the
 * group functions of xxhash.h) and a wave takes sixteen segments. */
#include "common.h"
#include "cudec_mi355x.h"
#include "xxhash.h"

namespace cudec_amd {
namespace {

constexpr int kBlock = 256;
constexpr uint32_t kSegmentsPerBlock = kBlock / xxh::kGroup;

/* XXH32 of segments i (sizes[i] bytes at ptrs[i]) into out[i]. */
__global__ void __launch_bounds__(kBlock)
    xxh32_batch(const void* const* ptrs, const size_t* sizes, size_t n, uint32_t seed,
                uint32_t* out) {
    const uint32_t lane = __lane_id();
    const uint32_t q = lane % xxh::kGroup;
    const size_t per_grid = (size_t)gridDim.x * kSegmentsPerBlock;
    for (size_t i = (size_t)blockIdx.x * kSegmentsPerBlock + threadIdx.x / xxh::kGroup;; i += per_grid) {
        /* every lane of the wave runs the loop the same number of times
         * (the group's shuffles need all four lanes) */
        const bool live = i < n;
        if (__ballot(live) == 0) {
            break;
        }
        const uint32_t h = xxh::xxh32_group(live ? (gcu8*)ptrs[i] : nullptr, live ? sizes[i] : 0,
                                            seed, q, lane - q);
        if (live && q == 0) {
            out[i] = h;
        }
    }
}

/* XXH64 of segments i into out[i]. */
__global__ void __launch_bounds__(kBlock)
    xxh64_batch(const void* const* ptrs, const size_t* sizes, size_t n, uint64_t seed,
                uint64_t* out) {
    const uint32_t lane = __lane_id();
    const uint32_t q = lane % xxh::kGroup;
    const size_t per_grid = (size_t)gridDim.x * kSegmentsPerBlock;
    for (size_t i = (size_t)blockIdx.x * kSegmentsPerBlock + threadIdx.x / xxh::kGroup;; i += per_grid) {
        const bool live = i < n;
        if (__ballot(live) == 0) {
            break;
        }
        const uint64_t h = xxh::xxh64_group(live ? (gcu8*)ptrs[i] : nullptr, live ? sizes[i] : 0,
                                            seed, q, lane - q);
        if (live && q == 0) {
            out[i] = h;
        }
    }
}

cudec_status checked_launch() {
    return hipGetLastError() == hipSuccess ? CUDEC_OK : CUDEC_ERR_CUDA;
}

unsigned grid_for(size_t n) {
    size_t g = (n + kSegmentsPerBlock - 1) / kSegmentsPerBlock;
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
