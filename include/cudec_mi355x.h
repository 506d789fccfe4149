/* MI355X-specific additions to the cudec C ABI. Nothing here changes
 * cudec.h; a caller of the reference library needs none of it.
 *
 * Launch forms. cudec_lz4_decompress_batch and cudec_snappy_decompress_batch
 * run the tuned form. The *_plain entries run the plain form (one wave64
 * per chunk, grid-stride, no persistence), kept as the drop-in baseline;
 * their contract is exactly the batch contract of cudec.h. Both forms
 * produce identical bytes, statuses and bytes_written for every input. */
#ifndef CUDEC_MI355X_H
#define CUDEC_MI355X_H

#include "cudec.h"

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* No counterpart in the reference: the plain launch form of
 * cudec_lz4_decompress_batch (include/cudec.h:102-108), for comparison. */
cudec_status cudec_lz4_decompress_batch_plain(const void* const* d_src_ptrs,
                                              const size_t* d_src_sizes,
                                              void* const* d_dst_ptrs,
                                              const size_t* d_dst_capacities,
                                              size_t chunk_count,
                                              cudec_chunk_result* d_results,
                                              cudec_stream_t stream);

/* No counterpart in the reference: the plain launch form of
 * cudec_snappy_decompress_batch. */
cudec_status cudec_snappy_decompress_batch_plain(
    const void* const* d_src_ptrs, const size_t* d_src_sizes,
    void* const* d_dst_ptrs, const size_t* d_dst_capacities,
    size_t chunk_count, cudec_chunk_result* d_results, cudec_stream_t stream);

/* The largest chunk_count the batch entries accept (restates
 * src/batch_limits.h:28 for wave64: INT32_MAX * 64). */
size_t cudec_mi355x_max_batch_chunks(void);

/* No counterpart in the reference: launch tuning of the tuned batch
 * entries (cudec_lz4_decompress_batch, cudec_snappy_decompress_batch) on
 * the current device. A chunk's cost is capacity + (capacity - compressed
 * size): the output, plus what matches add to it. Costs fall into
 * CUDEC_MI355X_COST_BINS bins, four per octave, bin 0 the costliest
 * (cudec_mi355x_cost_bin). A format's decode plan names, for each bin, the
 * lanes that decode one of its chunks: 128 (two waves of a block, one
 * parsing the next window while the other executes the current one) or 64
 * (one wave). Inside one launch the two-wave chunks start first, then the
 * one-wave ones, each costliest first. Output bytes and results never
 * depend on the plan. Plans are per device and per format, process-wide;
 * the default puts chunks costing at least 512 KiB on two waves. */
#define CUDEC_MI355X_COST_BINS 256
#define CUDEC_MI355X_FORMAT_LZ4 0
#define CUDEC_MI355X_FORMAT_SNAPPY 1

/* The cost bin of a chunk. */
int cudec_mi355x_cost_bin(size_t compressed_size, size_t capacity);

/* Sets or reads a format's plan: h_lanes_per_bin holds
 * CUDEC_MI355X_COST_BINS entries, each 64 or 128 (anything else is
 * CUDEC_ERR_INVALID_ARGUMENT, as are an unknown format and NULL). */
cudec_status cudec_mi355x_set_decode_plan(int format, const uint16_t* h_lanes_per_bin);
cudec_status cudec_mi355x_decode_plan(int format, uint16_t* h_lanes_per_bin);

/* A format's plan by threshold: chunks whose cost is at least min_cost
 * bytes decode on two waves, cheaper ones on one; 0 puts every chunk on
 * two waves, SIZE_MAX none. The getters return the threshold last set
 * (a plan set per bin since then is read with cudec_mi355x_decode_plan).
 * The LZ4 setting applies to cudec_lz4_decompress_batch only, the Snappy
 * one to cudec_snappy_decompress_batch only. */
cudec_status cudec_mi355x_set_lz4_pipeline_threshold(size_t min_cost);
cudec_status cudec_mi355x_lz4_pipeline_threshold(size_t* min_cost);
cudec_status cudec_mi355x_set_snappy_pipeline_threshold(size_t min_cost);
cudec_status cudec_mi355x_snappy_pipeline_threshold(size_t* min_cost);

/* No counterpart in the reference: the autotuner of the LZ4 threshold. It
 * decodes the caller's batch (arguments as cudec_lz4_decompress_batch) once
 * per candidate threshold after a warmup, synchronously, keeps the fastest
 * for the current device and returns it in *h_best_min_cost. With them
 * non-NULL, h_candidates and h_ms (CUDEC_MI355X_TUNE_CANDIDATES entries)
 * receive each candidate and its best event-timed launch in milliseconds.
 * The batch's outputs and results are those of the last run. */
#define CUDEC_MI355X_TUNE_CANDIDATES 9
cudec_status cudec_mi355x_autotune_lz4(const void* const* d_src_ptrs,
                                       const size_t* d_src_sizes, void* const* d_dst_ptrs,
                                       const size_t* d_dst_capacities, size_t chunk_count,
                                       cudec_chunk_result* d_results, cudec_stream_t stream,
                                       size_t* h_best_min_cost, size_t* h_candidates,
                                       float* h_ms);

/* No counterpart in the reference: the per-bin autotuner of a format's
 * plan, on the caller's batch (arguments as the format's batch entry),
 * synchronously. It first sweeps the thresholds above, then, from the
 * fastest threshold plan, visits each octave of cost the batch occupies,
 * costliest first, and moves that octave's bins to the other lane count;
 * a move is kept when the batch then decodes at least 1% faster (best of
 * two event-timed launches after a warmup). The final plan is set for the
 * current device and written to h_lanes_per_bin (CUDEC_MI355X_COST_BINS
 * entries). With them non-NULL, *h_ms_threshold receives the best
 * threshold plan's time and *h_ms_plan the final plan's, in milliseconds.
 * The batch's outputs and results are those of the last run. */
cudec_status cudec_mi355x_autotune_plan(int format, const void* const* d_src_ptrs,
                                        const size_t* d_src_sizes, void* const* d_dst_ptrs,
                                        const size_t* d_dst_capacities, size_t chunk_count,
                                        cudec_chunk_result* d_results, cudec_stream_t stream,
                                        uint16_t* h_lanes_per_bin, float* h_ms_threshold,
                                        float* h_ms_plan);

/* No counterpart in the reference. cudec_lz4_decompress_batch for callers
 * that cannot allocate device memory: code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc. Same arguments, contract,
 * statuses, bytes and results as cudec_lz4_decompress_batch (it runs the
 * same kernels), but the launch's scratch (its work queue) is a fixed
 * device array inside the library, about 4 MiB on each GPU it runs on, held
 * until the process ends, instead of a stream-ordered allocation. It calls
 * no allocating runtime routine. At most CUDEC_MI355X_NOALLOC_MAX_CHUNKS
 * chunks per call: more is CUDEC_ERR_INVALID_ARGUMENT, with no device call.
 * Every call on one GPU uses that GPU's one array: calls whose launches
 * overlap in time on one GPU (different streams, other host threads) race
 * and may decode wrongly, so order them (one stream, or events). */
#define CUDEC_MI355X_NOALLOC_MAX_CHUNKS 1048576
cudec_status cudec_mi355x_lz4_decompress_batch_noalloc(const void* const* d_src_ptrs,
                                                       const size_t* d_src_sizes,
                                                       void* const* d_dst_ptrs,
                                                       const size_t* d_dst_capacities,
                                                       size_t chunk_count,
                                                       cudec_chunk_result* d_results,
                                                       cudec_stream_t stream);

/* The parse ceiling, for measurement (the reference keeps the same arm:
 * lz4_decode_batch<ParseOnly>, timed beside the full decode by its GPU
 * bench). The format's tuned batch launch (format: CUDEC_MI355X_FORMAT_*)
 * with every window parsed and every status decided, and nothing copied:
 * d_results receives the same statuses and bytes_written as the format's
 * batch entry, and no destination byte is written. Arguments and
 * validation as the batch entry; at most 2^32 - 1 chunks. */
cudec_status cudec_mi355x_parse_batch(int format, const void* const* d_src_ptrs,
                                      const size_t* d_src_sizes, void* const* d_dst_ptrs,
                                      const size_t* d_dst_capacities, size_t chunk_count,
                                      cudec_chunk_result* d_results, cudec_stream_t stream);

/* Device checksums (the reference computes them on the host,
 * src/xxhash32.h and src/xxhash64.h): XXH32 and XXH64 with the given seed,
 * as the xxHash specification defines them, of count segments: segment i
 * is d_sizes[i] bytes at d_ptrs[i], and its hash lands in d_hashes[i].
 * The LZ4 frame format checksums with XXH32 (seed 0), the Zstd frame
 * format's content checksum is the low 32 bits of XXH64 (seed 0). Arrays
 * are device memory; asynchronous on stream, like the batch entries.
 * NULL arrays or count 0 are CUDEC_ERR_INVALID_ARGUMENT. */
cudec_status cudec_mi355x_xxh32_batch(const void* const* d_ptrs, const size_t* d_sizes,
                                      size_t count, uint32_t seed, uint32_t* d_hashes,
                                      cudec_stream_t stream);
cudec_status cudec_mi355x_xxh64_batch(const void* const* d_ptrs, const size_t* d_sizes,
                                      size_t count, uint64_t seed, uint64_t* d_hashes,
                                      cudec_stream_t stream);

/* No counterpart in the reference: a 64-bit digest of each chunk's bytes,
 * computed on the device, for checking decoded output in place.
 * Chunk i covers d_ptrs[i][0, len_i), where len_i is d_results[i]
 * .bytes_written when d_results is given (0 unless its status is OK),
 * else d_sizes[i]. Definition: len * 0x9E3779B97F4A7C15 plus, over the
 * 8-byte little-endian words w_j of the zero-padded bytes, the sum of
 * splitmix64_finalise(w_j + j * 0xD6E8FEB86659FD93), all mod 2^64.
 * Arrays are device memory; asynchronous on stream like the batch entry. */
cudec_status cudec_mi355x_digest_batch(const void* const* d_ptrs, const size_t* d_sizes,
                                       const cudec_chunk_result* d_results,
                                       size_t chunk_count, uint64_t* d_digests,
                                       cudec_stream_t stream);

/* ---------------------------------------------------------------- serving
 *
 * No counterpart in the reference. A request loop over the tuned LZ4
 * batch decoder, run in real time on the calling thread:
 *
 *  - request r arrives at h_requests[r].arrival_s seconds after the call
 *    starts; its chunks are entries [first_chunk, first_chunk + chunk_count)
 *    of the chunk tables, whose compressed bytes are already in device
 *    memory (h_src_ptrs holds device addresses, in a host array);
 *  - admission is first come, first served against an HBM budget: a
 *    request is admitted once its output (the sum of its chunks'
 *    capacities, each rounded up to 256 bytes) fits in the free part of a
 *    device output pool of hbm_budget bytes; the pool is allocated once;
 *  - continuous batching: every admitted request not yet launched goes into
 *    the next launch, on any of max_inflight streams that is idle, while
 *    earlier launches are still running; a request's output is released
 *    back to the pool when its launch completes;
 *  - per chunk, the result record lands in h_results, and with h_digests
 *    given, the digest of its output (cudec_mi355x_digest_batch) in
 *    h_digests; per request, h_latency_s gets completion time minus
 *    scheduled arrival (completion = when the loop sees its launch done).
 *
 * Returns CUDEC_ERR_INVALID_ARGUMENT (synchronously, before any device
 * call) for NULL arrays, request_count or chunk_count 0, a request range
 * outside the chunk tables, arrivals out of order, a request larger than
 * the budget, or max_inflight outside [1, 16]; CUDEC_ERR_CUDA for a device
 * failure; else CUDEC_OK (per-chunk outcomes are in h_results). */
typedef struct cudec_mi355x_request {
    double arrival_s;
    size_t first_chunk;
    size_t chunk_count;
} cudec_mi355x_request;

typedef struct cudec_mi355x_serve_stats {
    double elapsed_s;          /* first arrival to last completion */
    uint64_t launches;         /* batch launches */
    uint64_t max_inflight;     /* most launches in flight at once */
    uint64_t max_batch_requests; /* most requests in one launch */
    uint64_t peak_pool_bytes;  /* most pool bytes held at once */
    uint64_t output_bytes;     /* sum of bytes_written */
    uint64_t compressed_bytes; /* sum of compressed sizes */
} cudec_mi355x_serve_stats;

cudec_status cudec_mi355x_serve_lz4(const void* const* h_src_ptrs, const size_t* h_src_sizes,
                                    const size_t* h_dst_caps, size_t chunk_count,
                                    const cudec_mi355x_request* h_requests,
                                    size_t request_count, size_t hbm_budget,
                                    int max_inflight, cudec_chunk_result* h_results,
                                    uint64_t* h_digests, double* h_latency_s,
                                    cudec_mi355x_serve_stats* stats);

#ifdef __cplusplus
}
#endif

#endif /* CUDEC_MI355X_H */
