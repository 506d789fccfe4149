/* The C ABI entry points of the batch decoders.
 *
 * Each entry validates synchronously with no device-runtime call, consumes
 * the thread's pending runtime error, submits one launch on the caller's
 * stream, and reports only submission (src/batch.cu:6-31,
 * src/lz4_decode.cuh:187-201). No exception crosses this boundary. */
#include "common.h"

#include "cudec_mi355x.h"
#include "cudec_snappy.h"
#include "cudec_zstd.h"

namespace cudec_amd {

static_assert(kCostBins == CUDEC_MI355X_COST_BINS, "one bin count");
static_assert(kFormatLz4 == CUDEC_MI355X_FORMAT_LZ4 && kFormatSnappy == CUDEC_MI355X_FORMAT_SNAPPY,
              "one format numbering");

/* Launch-form selector. */
enum Form : int { kFormPlain = 0, kFormTuned = 1, kFormTunedFixed = 2, kFormParseOnly = 3 };

/* parse_only: the parse-only build (statuses and sizes, no output bytes) */
hipError_t launch_tuned(int format, const void* const* srcs,
                        const size_t* src_sizes, void* const* dsts,
                        const size_t* caps, size_t n,
                        cudec_chunk_result* results, hipStream_t stream,
                        bool parse_only);
/* launch_tuned with the fixed per-device scratch array (n <=
 * CUDEC_MI355X_NOALLOC_MAX_CHUNKS): no allocation. */
hipError_t launch_tuned_fixed(int format, const void* const* srcs,
                              const size_t* src_sizes, void* const* dsts,
                              const size_t* caps, size_t n,
                              cudec_chunk_result* results, hipStream_t stream);

/* One wave per Zstandard frame (zstd.hip). */
hipError_t launch_zstd(const void* const* srcs, const size_t* sizes, void* const* dsts,
                       const size_t* caps, size_t n, cudec_chunk_result* results,
                       hipStream_t stream);

/* src/lz4_decode.cuh:187-201 with the wave64 bound. */
static cudec_status validate(const void* const* srcs, const size_t* sizes,
                             void* const* dsts, const size_t* caps, size_t n,
                             cudec_chunk_result* results) {
    if (srcs == nullptr || sizes == nullptr || dsts == nullptr ||
        caps == nullptr || results == nullptr ||
        reinterpret_cast<uintptr_t>(results) % 16 != 0 || n == 0 ||
        n > kMaxBatchChunks) {
        return CUDEC_ERR_INVALID_ARGUMENT;
    }
    return CUDEC_OK;
}

static cudec_status submit(int format, int form, const void* const* srcs,
                           const size_t* sizes, void* const* dsts,
                           const size_t* caps, size_t n,
                           cudec_chunk_result* results,
                           cudec_stream_t stream) noexcept {
    const cudec_status valid = validate(srcs, sizes, dsts, caps, n, results);
    if (valid != CUDEC_OK) {
        return valid;
    }
    if ((form == kFormTunedFixed && n > CUDEC_MI355X_NOALLOC_MAX_CHUNKS) ||
        (form == kFormParseOnly && n > UINT32_MAX)) {
        return CUDEC_ERR_INVALID_ARGUMENT;
    }
    (void)hipGetLastError();
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    const hipError_t err =
        form == kFormPlain ? launch_plain(format, srcs, sizes, dsts, caps, n, results, s)
        : form == kFormTunedFixed
            ? launch_tuned_fixed(format, srcs, sizes, dsts, caps, n, results, s)
            : launch_tuned(format, srcs, sizes, dsts, caps, n, results, s,
                           form == kFormParseOnly);
    return err == hipSuccess ? CUDEC_OK : CUDEC_ERR_CUDA;
}

}  // namespace cudec_amd

using namespace cudec_amd;

extern "C" {

int cudec_version(void) {
    return CUDEC_VERSION_MAJOR * 10000 + CUDEC_VERSION_MINOR * 100 +
           CUDEC_VERSION_PATCH;
}

cudec_status cudec_lz4_decompress_batch(const void* const* d_src_ptrs,
                                        const size_t* d_src_sizes,
                                        void* const* d_dst_ptrs,
                                        const size_t* d_dst_capacities,
                                        size_t chunk_count,
                                        cudec_chunk_result* d_results,
                                        cudec_stream_t stream) {
    return submit(kFormatLz4, kFormTuned, d_src_ptrs, d_src_sizes, d_dst_ptrs,
                  d_dst_capacities, chunk_count, d_results, stream);
}

cudec_status cudec_snappy_decompress_batch(const void* const* d_src_ptrs,
                                           const size_t* d_src_sizes,
                                           void* const* d_dst_ptrs,
                                           const size_t* d_dst_capacities,
                                           size_t chunk_count,
                                           cudec_chunk_result* d_results,
                                           cudec_stream_t stream) {
    return submit(kFormatSnappy, kFormTuned, d_src_ptrs, d_src_sizes,
                  d_dst_ptrs, d_dst_capacities, chunk_count, d_results,
                  stream);
}

cudec_status cudec_zstd_decompress_batch(const void* const* d_src_ptrs,
                                         const size_t* d_src_sizes,
                                         void* const* d_dst_ptrs,
                                         const size_t* d_dst_capacities,
                                         size_t chunk_count,
                                         cudec_chunk_result* d_results,
                                         cudec_stream_t stream) {
    const cudec_status valid = validate(d_src_ptrs, d_src_sizes, d_dst_ptrs, d_dst_capacities,
                                        chunk_count, d_results);
    if (valid != CUDEC_OK) {
        return valid;
    }
    if (chunk_count > UINT32_MAX) {
        return CUDEC_ERR_INVALID_ARGUMENT; /* frames are ordered by 32-bit index */
    }
    (void)hipGetLastError();
    const hipError_t err = launch_zstd(d_src_ptrs, d_src_sizes, d_dst_ptrs, d_dst_capacities,
                                       chunk_count, d_results,
                                       reinterpret_cast<hipStream_t>(stream));
    return err == hipSuccess ? CUDEC_OK : CUDEC_ERR_CUDA;
}

cudec_status cudec_lz4_decompress_batch_plain(const void* const* d_src_ptrs,
                                              const size_t* d_src_sizes,
                                              void* const* d_dst_ptrs,
                                              const size_t* d_dst_capacities,
                                              size_t chunk_count,
                                              cudec_chunk_result* d_results,
                                              cudec_stream_t stream) {
    return submit(kFormatLz4, kFormPlain, d_src_ptrs, d_src_sizes, d_dst_ptrs,
                  d_dst_capacities, chunk_count, d_results, stream);
}

cudec_status cudec_mi355x_lz4_decompress_batch_noalloc(
    const void* const* d_src_ptrs, const size_t* d_src_sizes,
    void* const* d_dst_ptrs, const size_t* d_dst_capacities,
    size_t chunk_count, cudec_chunk_result* d_results,
    cudec_stream_t stream) {
    return submit(kFormatLz4, kFormTunedFixed, d_src_ptrs, d_src_sizes,
                  d_dst_ptrs, d_dst_capacities, chunk_count, d_results,
                  stream);
}

cudec_status cudec_mi355x_parse_batch(int format, const void* const* d_src_ptrs,
                                      const size_t* d_src_sizes, void* const* d_dst_ptrs,
                                      const size_t* d_dst_capacities, size_t chunk_count,
                                      cudec_chunk_result* d_results, cudec_stream_t stream) {
    if (format != kFormatLz4 && format != kFormatSnappy) {
        return CUDEC_ERR_INVALID_ARGUMENT;
    }
    return submit(format, kFormParseOnly, d_src_ptrs, d_src_sizes, d_dst_ptrs,
                  d_dst_capacities, chunk_count, d_results, stream);
}

cudec_status cudec_snappy_decompress_batch_plain(
    const void* const* d_src_ptrs, const size_t* d_src_sizes,
    void* const* d_dst_ptrs, const size_t* d_dst_capacities,
    size_t chunk_count, cudec_chunk_result* d_results,
    cudec_stream_t stream) {
    return submit(kFormatSnappy, kFormPlain, d_src_ptrs, d_src_sizes,
                  d_dst_ptrs, d_dst_capacities, chunk_count, d_results,
                  stream);
}

static cudec_status set_threshold(int format, size_t min_cost) {
    (void)hipGetLastError();
    return set_pipe_min_cost(format, min_cost) == hipSuccess ? CUDEC_OK : CUDEC_ERR_CUDA;
}

static cudec_status get_threshold(int format, size_t* min_cost) {
    if (min_cost == nullptr) {
        return CUDEC_ERR_INVALID_ARGUMENT;
    }
    (void)hipGetLastError();
    uint64_t v = 0;
    if (get_pipe_min_cost(format, &v) != hipSuccess) {
        return CUDEC_ERR_CUDA;
    }
    *min_cost = (size_t)v;
    return CUDEC_OK;
}

cudec_status cudec_mi355x_set_lz4_pipeline_threshold(size_t min_cost) {
    return set_threshold(kFormatLz4, min_cost);
}

cudec_status cudec_mi355x_lz4_pipeline_threshold(size_t* min_cost) {
    return get_threshold(kFormatLz4, min_cost);
}

cudec_status cudec_mi355x_set_snappy_pipeline_threshold(size_t min_cost) {
    return set_threshold(kFormatSnappy, min_cost);
}

cudec_status cudec_mi355x_snappy_pipeline_threshold(size_t* min_cost) {
    return get_threshold(kFormatSnappy, min_cost);
}

int cudec_mi355x_cost_bin(size_t compressed_size, size_t capacity) {
    return (int)host_cost_bin(compressed_size, capacity);
}

cudec_status cudec_mi355x_set_decode_plan(int format, const uint16_t* h_lanes_per_bin) {
    if ((format != kFormatLz4 && format != kFormatSnappy) || h_lanes_per_bin == nullptr) {
        return CUDEC_ERR_INVALID_ARGUMENT;
    }
    for (int b = 0; b < kCostBins; b++) {
        if (h_lanes_per_bin[b] != 64 && h_lanes_per_bin[b] != 128) {
            return CUDEC_ERR_INVALID_ARGUMENT;
        }
    }
    (void)hipGetLastError();
    return set_decode_plan(format, h_lanes_per_bin) == hipSuccess ? CUDEC_OK : CUDEC_ERR_CUDA;
}

cudec_status cudec_mi355x_decode_plan(int format, uint16_t* h_lanes_per_bin) {
    if ((format != kFormatLz4 && format != kFormatSnappy) || h_lanes_per_bin == nullptr) {
        return CUDEC_ERR_INVALID_ARGUMENT;
    }
    (void)hipGetLastError();
    return get_decode_plan(format, h_lanes_per_bin) == hipSuccess ? CUDEC_OK : CUDEC_ERR_CUDA;
}

size_t cudec_mi355x_max_batch_chunks(void) { return kMaxBatchChunks; }

}  // extern "C"
