/* Additive Zstandard entry point for the cudec C ABI (this header does not
 * modify cudec.h).
 *
 * cudec_zstd_decompress_batch carries the cudec_lz4_decompress_batch
 * contract: the same device-side argument arrays, per-chunk results,
 * synchronous argument refusal, asynchronous launch on `stream`,
 * pending-error semantics, bytes_written 0 on any reject, and on success
 * writes touching exactly the bytes reported. Each chunk is one Zstandard
 * frame (RFC 8878) inside iderex/cudec's v1 decode subset
 * (docs/MASTERPLAN.md:1172-1305; decoded there only by host twins,
 * src/zstd_frame.h, src/zstd_literals.h, src/zstd_seq.h, src/zstd_repcode.h,
 * with no batch entry):
 *   - CUDEC_ERR_UNSUPPORTED for a skippable frame, a dictionary id, a frame
 *     with no content size, or a window above 8 MiB;
 *   - CUDEC_ERR_OUTPUT_TOO_SMALL when the declared content size, or the
 *     output, exceeds the chunk's capacity;
 *   - CUDEC_ERR_CORRUPT_INPUT for every other refusal: a malformed header,
 *     block, literals or sequences section, a bitstream with bits left
 *     over, a match reaching past the window or before the output, output
 *     other than the declared content size, a failed XXH64 content
 *     checksum, or bytes after the frame.
 * On success bytes_written is the declared content size. A chunk_count
 * above UINT32_MAX is refused with CUDEC_ERR_INVALID_ARGUMENT, like the
 * other argument errors, before any device call. */
#ifndef CUDEC_ZSTD_H
#define CUDEC_ZSTD_H

#include "cudec.h"

#ifdef __cplusplus
extern "C" {
#endif

cudec_status cudec_zstd_decompress_batch(const void* const* d_src_ptrs,
                                         const size_t* d_src_sizes,
                                         void* const* d_dst_ptrs,
                                         const size_t* d_dst_capacities,
                                         size_t chunk_count,
                                         cudec_chunk_result* d_results,
                                         cudec_stream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* CUDEC_ZSTD_H */
