/* Autotuners of the tuned batch launch: which chunks decode on two waves
 * (parse and execute pipelined) and which on one.
 *
 * Chunks fall into cost bins (cost = capacity + (capacity - compressed
 * size)); a format's decode plan names the lanes per chunk of every bin.
 * Two waves shorten one chunk's critical path (large chunks bound a
 * launch); one wave keeps twice as many chunks in flight (batches of small
 * chunks are throughput-bound). cudec_mi355x_autotune_lz4 sweeps threshold
 * plans; cudec_mi355x_autotune_plan refines the best threshold plan octave
 * by octave. Both decode the caller's batch per candidate after a warmup
 * and keep the fastest for the current device. */
#include <algorithm>
#include <cstring>
#include <vector>

#include "cudec_mi355x.h"
#include "cudec_snappy.h"
#include "hip_owner.h"

namespace cudec_amd {
namespace {

constexpr size_t kCandidates[CUDEC_MI355X_TUNE_CANDIDATES] = {
    0, 64u << 10, 128u << 10, 256u << 10, 512u << 10, 1u << 20, 2u << 20, 4u << 20, SIZE_MAX};
constexpr int kRepeats = 2;

using BatchEntry = cudec_status (*)(const void* const*, const size_t*, void* const*,
                                    const size_t*, size_t, cudec_chunk_result*, cudec_stream_t);

/* One batch, launched and timed on its stream. */
struct Timed {
    BatchEntry entry;
    const void* const* src;
    const size_t* sizes;
    void* const* dst;
    const size_t* caps;
    size_t n;
    cudec_chunk_result* results;
    cudec_stream_t stream;
    Event t0, t1;

    /* Best event-timed launch of kRepeats after one warmup, in *ms. */
    cudec_status best_of(float* ms) {
        const hipStream_t s = reinterpret_cast<hipStream_t>(stream);
        float fastest = -1.0f;
        for (int r = 0; r <= kRepeats; r++) { /* run 0 warms up */
            if (hipEventRecord(t0.get(), s) != hipSuccess) {
                return CUDEC_ERR_CUDA;
            }
            const cudec_status st = entry(src, sizes, dst, caps, n, results, stream);
            if (st != CUDEC_OK) {
                return st;
            }
            float t = 0.0f;
            if (hipEventRecord(t1.get(), s) != hipSuccess ||
                hipEventSynchronize(t1.get()) != hipSuccess ||
                hipEventElapsedTime(&t, t0.get(), t1.get()) != hipSuccess) {
                return CUDEC_ERR_CUDA;
            }
            if (r > 0 && (fastest < 0.0f || t < fastest)) {
                fastest = t;
            }
        }
        *ms = fastest;
        return CUDEC_OK;
    }
};

bool batch_args_valid(const void* const* src, const size_t* sizes, void* const* dst,
                      const size_t* caps, size_t n, const cudec_chunk_result* results) {
    return src != nullptr && sizes != nullptr && dst != nullptr && caps != nullptr &&
           results != nullptr && reinterpret_cast<uintptr_t>(results) % 16 == 0 && n != 0 &&
           n <= cudec_mi355x_max_batch_chunks();
}

}  // namespace
}  // namespace cudec_amd

extern "C" cudec_status cudec_mi355x_autotune_lz4(
    const void* const* d_src_ptrs, const size_t* d_src_sizes, void* const* d_dst_ptrs,
    const size_t* d_dst_capacities, size_t chunk_count, cudec_chunk_result* d_results,
    cudec_stream_t stream, size_t* h_best_min_cost, size_t* h_candidates, float* h_ms) {
    using namespace cudec_amd;
    if (!batch_args_valid(d_src_ptrs, d_src_sizes, d_dst_ptrs, d_dst_capacities, chunk_count,
                          d_results) ||
        h_best_min_cost == nullptr) {
        return CUDEC_ERR_INVALID_ARGUMENT;
    }
    try {
        size_t previous = 0;
        cudec_status st = cudec_mi355x_lz4_pipeline_threshold(&previous);
        if (st != CUDEC_OK) {
            return st;
        }
        const hipStream_t s = reinterpret_cast<hipStream_t>(stream);
        Event t0, t1;
        if (make_timing_event(t0) != hipSuccess || make_timing_event(t1) != hipSuccess) {
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
                if (hipEventRecord(t0.get(), s) != hipSuccess) {
                    return CUDEC_ERR_CUDA;
                }
                st = cudec_lz4_decompress_batch(d_src_ptrs, d_src_sizes, d_dst_ptrs,
                                                d_dst_capacities, chunk_count, d_results, stream);
                if (st != CUDEC_OK) {
                    (void)cudec_mi355x_set_lz4_pipeline_threshold(previous);
                    return st;
                }
                float ms = 0.0f;
                if (hipEventRecord(t1.get(), s) != hipSuccess ||
                    hipEventSynchronize(t1.get()) != hipSuccess ||
                    hipEventElapsedTime(&ms, t0.get(), t1.get()) != hipSuccess) {
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

extern "C" cudec_status cudec_mi355x_autotune_plan(
    int format, const void* const* d_src_ptrs, const size_t* d_src_sizes, void* const* d_dst_ptrs,
    const size_t* d_dst_capacities, size_t chunk_count, cudec_chunk_result* d_results,
    cudec_stream_t stream, uint16_t* h_lanes_per_bin, float* h_ms_threshold, float* h_ms_plan) {
    using namespace cudec_amd;
    if ((format != CUDEC_MI355X_FORMAT_LZ4 && format != CUDEC_MI355X_FORMAT_SNAPPY) ||
        !batch_args_valid(d_src_ptrs, d_src_sizes, d_dst_ptrs, d_dst_capacities, chunk_count,
                          d_results) ||
        h_lanes_per_bin == nullptr) {
        return CUDEC_ERR_INVALID_ARGUMENT;
    }
    const bool lz4 = format == CUDEC_MI355X_FORMAT_LZ4;
    auto set_threshold = lz4 ? cudec_mi355x_set_lz4_pipeline_threshold
                             : cudec_mi355x_set_snappy_pipeline_threshold;
    try {
        Timed run{lz4 ? cudec_lz4_decompress_batch : cudec_snappy_decompress_batch,
                  d_src_ptrs, d_src_sizes, d_dst_ptrs, d_dst_capacities, chunk_count, d_results,
                  stream, {}, {}};
        if (make_timing_event(run.t0) != hipSuccess || make_timing_event(run.t1) != hipSuccess) {
            return CUDEC_ERR_CUDA;
        }
        /* 1: the fastest threshold plan */
        float best = -1.0f;
        size_t best_cost = kCandidates[0];
        for (int c = 0; c < CUDEC_MI355X_TUNE_CANDIDATES; c++) {
            cudec_status st = set_threshold(kCandidates[c]);
            float ms = 0.0f;
            if (st == CUDEC_OK) {
                st = run.best_of(&ms);
            }
            if (st != CUDEC_OK) {
                return st;
            }
            if (best < 0.0f || ms < best) {
                best = ms;
                best_cost = kCandidates[c];
            }
        }
        cudec_status st = set_threshold(best_cost);
        if (st != CUDEC_OK) {
            return st;
        }
        if (h_ms_threshold != nullptr) {
            *h_ms_threshold = best;
        }
        std::vector<uint16_t> plan(CUDEC_MI355X_COST_BINS);
        st = cudec_mi355x_decode_plan(format, plan.data());
        if (st != CUDEC_OK) {
            return st;
        }
        /* 2: the octaves the batch occupies (bins 4k .. 4k + 3) */
        std::vector<size_t> sizes(chunk_count), caps(chunk_count);
        const hipStream_t s = reinterpret_cast<hipStream_t>(stream);
        if (hipMemcpyAsync(sizes.data(), d_src_sizes, chunk_count * sizeof(size_t),
                           hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipMemcpyAsync(caps.data(), d_dst_capacities, chunk_count * sizeof(size_t),
                           hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess) {
            return CUDEC_ERR_CUDA;
        }
        constexpr int kPerOctave = 4;
        std::vector<bool> used(CUDEC_MI355X_COST_BINS, false);
        for (size_t i = 0; i < chunk_count; i++) {
            used[(size_t)cudec_mi355x_cost_bin(sizes[i], caps[i])] = true;
        }
        /* 3: move each occupied octave to the other lane count, costliest
         * octave first; keep a move that gains at least 1% */
        for (int o = 0; o < CUDEC_MI355X_COST_BINS / kPerOctave; o++) {
            std::vector<uint16_t> trial = plan;
            bool any = false;
            bool all_two = true;
            for (int b = o * kPerOctave; b < (o + 1) * kPerOctave; b++) {
                if (used[(size_t)b]) {
                    any = true;
                    all_two = all_two && plan[(size_t)b] == 128;
                }
            }
            if (!any) {
                continue;
            }
            for (int b = o * kPerOctave; b < (o + 1) * kPerOctave; b++) {
                if (used[(size_t)b]) {
                    trial[(size_t)b] = all_two ? 64 : 128;
                }
            }
            float ms = 0.0f;
            st = cudec_mi355x_set_decode_plan(format, trial.data());
            if (st == CUDEC_OK) {
                st = run.best_of(&ms);
            }
            if (st != CUDEC_OK) {
                return st;
            }
            if (ms < 0.99f * best) {
                best = ms;
                plan = trial;
            }
        }
        st = cudec_mi355x_set_decode_plan(format, plan.data());
        if (st != CUDEC_OK) {
            return st;
        }
        std::copy(plan.begin(), plan.end(), h_lanes_per_bin);
        if (h_ms_plan != nullptr) {
            *h_ms_plan = best;
        }
        return CUDEC_OK;
    } catch (...) {
        return CUDEC_ERR_CUDA;
    }
}
