/* Additive Snappy entry point for the cudec C ABI (this header does not
 * modify cudec.h).
 *
 * cudec_snappy_decompress_batch carries the cudec_lz4_decompress_batch
 * contract with no deltas, as iderex/cudec commits to for its Snappy batch
 * entry (docs/MASTERPLAN.md:722-733; the entry is not shipped there): the
 * same device-side argument arrays, per-chunk results, synchronous
 * argument refusal, asynchronous launch on `stream`, pending-error
 * semantics, bytes_written 0 on any reject, and writes touching exactly the
 * bytes reported. Each chunk is one raw Snappy block (snappy::RawCompress
 * output). Its declared length is checked against the chunk's capacity
 * (CUDEC_ERR_OUTPUT_TOO_SMALL when it exceeds it); every other refusal is
 * CUDEC_ERR_CORRUPT_INPUT; on success bytes_written is the declared length.
 */
#ifndef CUDEC_SNAPPY_H
#define CUDEC_SNAPPY_H

#include "cudec.h"

#ifdef __cplusplus
extern "C" {
#endif

cudec_status cudec_snappy_decompress_batch(const void* const* d_src_ptrs,
                                           const size_t* d_src_sizes,
                                           void* const* d_dst_ptrs,
                                           const size_t* d_dst_capacities,
                                           size_t chunk_count,
                                           cudec_chunk_result* d_results,
                                           cudec_stream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* CUDEC_SNAPPY_H */
