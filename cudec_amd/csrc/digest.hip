/* Per-chunk output digest on the device: a 64-bit order-sensitive sum that
 * lets a caller check decoded chunks without copying them to the host.
 *
 *   digest(b[0, n)) = n * kLenMul + sum over 8-byte words j of
 *                     mix64(word_j + j * kPosMul)   (mod 2^64)
 *
 * word_j is bytes [8j, 8j + 8) little-endian, zero-padded past n; mix64 is
 * the SplitMix64 finaliser. The sum commutes, so lanes reduce in any order
 * and the result is still a pure function of the bytes. Not a
 * cryptographic hash: it detects corrupted or misplaced output. */
#include "common.h"
#include "cudec_mi355x.h"

namespace cudec_amd {
namespace {

constexpr uint64_t kLenMul = 0x9E3779B97F4A7C15ull;
constexpr uint64_t kPosMul = 0xD6E8FEB86659FD93ull;

__device__ __forceinline__ uint64_t mix64(uint64_t x) {
    x ^= x >> 30;
    x *= 0xBF58476D1CE4E5B9ull;
    x ^= x >> 27;
    x *= 0x94D049BB133111EBull;
    x ^= x >> 31;
    return x;
}

/* One wave per chunk, grid-stride. Each lane folds 16-byte pieces (two
 * words) of the chunk; a partial piece at the end is read bytewise. With
 * results given, a chunk's length is its bytes_written (0 unless OK). */
__global__ void __launch_bounds__(256) digest_batch(const void* const* ptrs, const size_t* sizes,
                                                    const cudec_chunk_result* results, size_t n,
                                                    uint64_t* out) {
    const uint32_t lane = __lane_id();
    const size_t waves = (size_t)gridDim.x * (blockDim.x / kWave);
    for (size_t c = (size_t)blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave; c < n;
         c += waves) {
        const uint8_t* p = uniform_ptr(static_cast<const uint8_t*>(ptrs[c]));
        uint64_t len;
        if (results != nullptr) {
            const cudec_chunk_result r = results[c];
            len = r.status == CUDEC_OK ? r.bytes_written : 0;
        } else {
            len = sizes[c];
        }
        len = uniform64(len);
        uint64_t acc = 0;
        const uint64_t whole = len / 16;
        for (uint64_t k = lane; k < whole; k += kWave) {
            uint64_t w[2];
            __builtin_memcpy(w, p + 16 * k, 16);
            acc += mix64(w[0] + (2 * k) * kPosMul) + mix64(w[1] + (2 * k + 1) * kPosMul);
        }
        const uint64_t rest = len - 16 * whole;
        if (rest != 0 && lane < 2) {
            /* the last one or two words, zero-padded */
            const uint64_t j = 2 * whole + lane;
            const uint64_t lo = 8 * j;
            if (lo < len) {
                uint64_t w = 0;
                const uint64_t m = len - lo < 8 ? len - lo : 8;
                for (uint64_t i = 0; i < m; i++) {
                    w |= (uint64_t)p[lo + i] << (8 * i);
                }
                acc += mix64(w + j * kPosMul);
            }
        }
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) {
            const uint32_t lo32 = (uint32_t)__shfl_xor((int)(uint32_t)acc, d, kWave);
            const uint32_t hi32 = (uint32_t)__shfl_xor((int)(uint32_t)(acc >> 32), d, kWave);
            acc += ((uint64_t)hi32 << 32) | lo32;
        }
        /* every lane holds the same total; all store it (no lane-0 branch) */
        out[c] = len * kLenMul + acc;
    }
}

}  // namespace
}  // namespace cudec_amd

extern "C" cudec_status cudec_mi355x_digest_batch(const void* const* d_ptrs,
                                                  const size_t* d_sizes,
                                                  const cudec_chunk_result* d_results,
                                                  size_t chunk_count, uint64_t* d_digests,
                                                  cudec_stream_t stream) {
    using namespace cudec_amd;
    if (d_ptrs == nullptr || d_digests == nullptr || (d_sizes == nullptr && d_results == nullptr) ||
        chunk_count == 0 || chunk_count > kMaxBatchChunks) {
        return CUDEC_ERR_INVALID_ARGUMENT;
    }
    (void)hipGetLastError();
    const size_t per_block = 256 / kWave;
    size_t blocks = (chunk_count + per_block - 1) / per_block;
    if (blocks > 8192) {
        blocks = 8192;
    }
    hipLaunchKernelGGL(digest_batch, dim3((unsigned)blocks), dim3(256), 0,
                       reinterpret_cast<hipStream_t>(stream), d_ptrs, d_sizes, d_results,
                       chunk_count, d_digests);
    return hipGetLastError() == hipSuccess ? CUDEC_OK : CUDEC_ERR_CUDA;
}
