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

#ifdef __cplusplus
}
#endif

#endif /* CUDEC_MI355X_H */
