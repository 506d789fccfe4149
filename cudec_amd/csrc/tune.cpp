/* Autotuner of the LZ4 batch launch: which chunks decode on two waves
 * (parse and execute pipelined) and which on one.
 *
 * Chunks are ordered by cost (capacity + (capacity - compressed size))
 * inside a launch; a chunk at or above the threshold cost runs on a
 * two-wave block, a
 * cheaper one on a single wave. Two waves shorten one chunk's critical path
 * (large chunks bound a launch); one wave keeps twice as many chunks in
 * flight (batches of small chunks are throughput-bound). The tuner decodes
 * the caller's batch once per candidate threshold after a warmup, keeps the
 * fastest for the current device, and reports every time. */
#include <cstring>

#include "cudec_mi355x.h"
#include "hip_owner.h"

namespace cudec_amd {
namespace {

constexpr size_t kCandidates[CUDEC_MI355X_TUNE_CANDIDATES] = {
    0, 64u << 10, 128u << 10, 256u << 10, 512u << 10, 1u << 20, 2u << 20, 4u << 20, SIZE_MAX};
constexpr int kRepeats = 2;

}  // namespace
}  // namespace cudec_amd

extern "C" cudec_status cudec_mi355x_autotune_lz4(
    const void* const* d_src_ptrs, const size_t* d_src_sizes, void* const* d_dst_ptrs,
    const size_t* d_dst_capacities, size_t chunk_count, cudec_chunk_result* d_results,
    cudec_stream_t stream, size_t* h_best_min_cost, size_t* h_candidates, float* h_ms) {
    using namespace cudec_amd;
    if (d_src_ptrs == nullptr || d_src_sizes == nullptr || d_dst_ptrs == nullptr ||
        d_dst_capacities == nullptr || d_results == nullptr ||
        reinterpret_cast<uintptr_t>(d_results) % 16 != 0 || chunk_count == 0 ||
        chunk_count > cudec_mi355x_max_batch_chunks() || h_best_min_cost == nullptr) {
        return CUDEC_ERR_INVALID_ARGUMENT;
    }
    try {
        size_t previous = 0;
        cudec_status st = cudec_mi355x_lz4_pipeline_threshold(&previous);
        if (st != CUDEC_OK) {
            return st;
        }
        const hipStream_t s = reinterpret_cast<hipStream_t>(stream);
        EventOwner t0, t1;
        if (hipEventCreate(&t0.e) != hipSuccess || hipEventCreate(&t1.e) != hipSuccess) {
            return CUDEC_ERR_CUDA;
        }
        float best = -1.0f;
        size_t best_cost = previous;
        for (int c = 0; c < CUDEC_MI355X_TUNE_CANDIDATES; c++) {
            st = cudec_mi355x_set_lz4_pipeline_threshold(kCandidates[c]);
            if (st != CUDEC_OK) {
                return st;
            }
            float fastest = -1.0f;
            for (int r = 0; r <= kRepeats; r++) { /* run 0 warms up */
                if (hipEventRecord(t0.e, s) != hipSuccess) {
                    return CUDEC_ERR_CUDA;
                }
                st = cudec_lz4_decompress_batch(d_src_ptrs, d_src_sizes, d_dst_ptrs,
                                                d_dst_capacities, chunk_count, d_results, stream);
                if (st != CUDEC_OK) {
                    (void)cudec_mi355x_set_lz4_pipeline_threshold(previous);
                    return st;
                }
                float ms = 0.0f;
                if (hipEventRecord(t1.e, s) != hipSuccess ||
                    hipEventSynchronize(t1.e) != hipSuccess ||
                    hipEventElapsedTime(&ms, t0.e, t1.e) != hipSuccess) {
                    (void)cudec_mi355x_set_lz4_pipeline_threshold(previous);
                    return CUDEC_ERR_CUDA;
                }
                if (r > 0 && (fastest < 0.0f || ms < fastest)) {
                    fastest = ms;
                }
            }
            if (h_candidates != nullptr) {
                h_candidates[c] = kCandidates[c];
            }
            if (h_ms != nullptr) {
                h_ms[c] = fastest;
            }
            if (best < 0.0f || fastest < best) {
                best = fastest;
                best_cost = kCandidates[c];
            }
        }
        st = cudec_mi355x_set_lz4_pipeline_threshold(best_cost);
        if (st != CUDEC_OK) {
            return st;
        }
        *h_best_min_cost = best_cost;
        return CUDEC_OK;
    } catch (...) {
        return CUDEC_ERR_CUDA;
    }
}
