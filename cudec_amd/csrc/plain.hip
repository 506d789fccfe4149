/* The plain launch form: one wave64 per chunk, grid-stride, no persistence.
 *
 * Every lane of a wave holds the same cursor (pinned with readfirstlane, so
 * the ladder runs on the scalar unit); the literal copy fans out one byte per
 * lane, and a match is the closed-form modular gather
 * dst[m + i] = dst[m - off + (i mod off)], with a wavefront-ordered fence
 * between the two copies and after the match. This is the wave64
 * counterpart of iderex/cudec's warp-per-chunk kernel
 * (src/lz4_decode.cuh:29-173) and stays beside the tuned form as the
 * drop-in baseline; Snappy elements run through the same body. */
#include "common.h"
#include "ladder.h"

namespace cudec_amd {

constexpr unsigned kPlainBlock = 256;           /* 4 waves */
constexpr unsigned kPlainMaxBlocks = 8192;

/* A literal run followed by a match, in chunk-relative offsets. */
__device__ __forceinline__ void plain_execute(gcu8* src, gu8* dst,
                                              uint32_t lane, uint64_t lit_src,
                                              uint64_t lit_dst,
                                              uint64_t lit_len,
                                              uint64_t match_dst,
                                              uint64_t match_len,
                                              uint64_t offset) {
    for (uint64_t i = lane; i < lit_len; i += kWave) {
        dst[lit_dst + i] = src[lit_src + i];
    }
    wave_order();
    if (match_len != 0) {
        const uint64_t from = match_dst - offset;
        if (match_len <= UINT32_MAX && offset <= UINT32_MAX) {
            const uint32_t off32 = (uint32_t)offset;
            for (uint64_t i = lane; i < match_len; i += kWave) {
                dst[match_dst + i] = dst[from + ((uint32_t)i % off32)];
            }
        } else {
            for (uint64_t i = lane; i < match_len; i += kWave) {
                dst[match_dst + i] = dst[from + (i % offset)];
            }
        }
        wave_order();
    }
}

template <int kFormat>
__global__ void __launch_bounds__(kPlainBlock)
    plain_batch(const void* const* srcs, const size_t* src_sizes,
                void* const* dsts, const size_t* caps, size_t n,
                cudec_chunk_result* results) {
    /* Geometry refusal: a block that is not a whole number of waves would
     * split a chunk's wave; the grid-uniform test strands no lane. */
    if (blockDim.x % kWave != 0 || gridDim.x == 0) {
        return;
    }
    const uint32_t lane = __lane_id();
    const uint64_t waves_per_block = blockDim.x / kWave;
    const uint64_t total_waves = (uint64_t)gridDim.x * waves_per_block;
    uint64_t chunk = uniform64((uint64_t)blockIdx.x * waves_per_block +
                               threadIdx.x / kWave);
    for (; chunk < n; chunk += total_waves) {
        gcu8* src = global_bytes(srcs[chunk]);
        gu8* dst = global_bytes(dsts[chunk]);
        const uint64_t src_size = uniform64(src_sizes[chunk]);
        const uint64_t cap = uniform64(caps[chunk]);
        /* Fuel: every step consumes at least one source byte, so
         * src_size + 1 steps (saturating) are never reached by a live
         * parse; exhausting it is a reject (src/lz4_decode.cuh:90-99). */
        uint64_t fuel = src_size == UINT64_MAX ? UINT64_MAX : src_size + 1;
        int status = kOk;
        bool done = false;
        uint64_t produced = 0;
        if (kFormat == kFormatLz4) {
            Lz4Cursor c{src, src_size, cap, 0, 0};
            Lz4Step q;
            for (;;) {
                status = lz4_next(c, q, done);
                if (status != kOk) {
                    break;
                }
                plain_execute(src, dst, lane, q.lit_src, q.lit_dst, q.lit_len,
                              q.match_dst, q.match_len, q.offset);
                if (done || fuel-- == 0) {
                    break;
                }
            }
            produced = c.dp;
        } else {
            SnappyCursor c{src, src_size, cap, 0, 0, 0};
            SnappyStep e;
            status = snappy_begin(c);
            while (status == kOk) {
                status = snappy_next(c, e, done);
                if (status != kOk) {
                    break;
                }
                if (e.is_copy) {
                    plain_execute(src, dst, lane, 0, e.to, 0, e.to, e.length,
                                  e.offset);
                } else {
                    plain_execute(src, dst, lane, e.from, e.to, e.length,
                                  e.to + e.length, 0, 1);
                }
                if (done || fuel-- == 0) {
                    break;
                }
            }
            produced = c.dp;
        }
        if (status == kOk && !done) {
            status = kCorrupt;
        }
        if (lane == 0) {
            write_result(results, chunk, status, produced);
        }
    }
}

hipError_t launch_plain(int format, const void* const* srcs,
                        const size_t* src_sizes, void* const* dsts,
                        const size_t* caps, size_t n,
                        cudec_chunk_result* results, hipStream_t stream) {
    const size_t per_block = kPlainBlock / kWave;
    size_t blocks = (n + per_block - 1) / per_block;
    if (blocks > kPlainMaxBlocks) {
        blocks = kPlainMaxBlocks;
    }
    if (format == kFormatSnappy) {
        hipLaunchKernelGGL(plain_batch<kFormatSnappy>, dim3((unsigned)blocks),
                           dim3(kPlainBlock), 0, stream, srcs, src_sizes, dsts,
                           caps, n, results);
    } else {
        hipLaunchKernelGGL(plain_batch<kFormatLz4>, dim3((unsigned)blocks),
                           dim3(kPlainBlock), 0, stream, srcs, src_sizes, dsts,
                           caps, n, results);
    }
    return hipGetLastError();
}

}  // namespace cudec_amd
