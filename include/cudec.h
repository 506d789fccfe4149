/* cudec public C ABI, as implemented by the MI355X (gfx950) library.
 *
 * Every declaration below is binary-identical to iderex/cudec's
 * include/cudec.h (function names and signatures, enum names and values,
 * struct layout), so a program built against that header links against this
 * library unchanged. The comments are this implementation's own account of
 * the contract it honours.
 *
 * Plain C: no HIP or CUDA header is needed to include it. No entry point
 * throws across this boundary, and no output is reported valid unless it
 * was fully validated.
 */
#ifndef CUDEC_H
#define CUDEC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Library version; cudec_version() returns
 * MAJOR * 10000 + MINOR * 100 + PATCH. */
#define CUDEC_VERSION_MAJOR 0
#define CUDEC_VERSION_MINOR 1
#define CUDEC_VERSION_PATCH 0

int cudec_version(void);

typedef enum cudec_status {
    CUDEC_OK = 0,
    CUDEC_ERR_INVALID_ARGUMENT = 1,
    CUDEC_ERR_CORRUPT_INPUT = 2,
    CUDEC_ERR_OUTPUT_TOO_SMALL = 3,
    /* A device-runtime failure (named for the original runtime; on this
     * library it reports a HIP error). */
    CUDEC_ERR_CUDA = 4,
    /* Reserved; no entry point returns it. */
    CUDEC_ERR_NOT_IMPLEMENTED = 5,
    /* A valid LZ4 frame outside the decoded subset (linked blocks or a
     * dictionary id). */
    CUDEC_ERR_UNSUPPORTED = 6
} cudec_status;

/* An opaque stream handle. On this library it is a hipStream_t passed
 * through unchanged; NULL selects the null stream. */
typedef struct CUstream_st* cudec_stream_t;

/* Per-chunk outcome written by the device. Fixed 16-byte layout. */
typedef struct cudec_chunk_result {
    int32_t status;         /* a cudec_status value */
    uint32_t reserved;      /* always written as 0 */
    uint64_t bytes_written; /* decoded bytes on CUDEC_OK, else 0 */
} cudec_chunk_result;

#if defined(__cplusplus) && __cplusplus >= 201103L
static_assert(sizeof(cudec_chunk_result) == 16 &&
                  offsetof(cudec_chunk_result, bytes_written) == 8,
              "cudec_chunk_result is a fixed 16-byte record");
#elif defined(__STDC_VERSION__) && __STDC_VERSION__ >= 201112L
_Static_assert(sizeof(cudec_chunk_result) == 16 &&
                   offsetof(cudec_chunk_result, bytes_written) == 8,
               "cudec_chunk_result is a fixed 16-byte record");
#else
typedef char cudec_chunk_result_layout_check
    [(sizeof(cudec_chunk_result) == 16 &&
      offsetof(cudec_chunk_result, bytes_written) == 8)
         ? 1
         : -1];
#endif

/* Batch LZ4 block decode.
 *
 * Inputs are device arrays of chunk_count entries; the pointers they hold
 * are device pointers of any alignment. d_results must be 16-byte aligned.
 * Asynchronous on `stream`: the return value covers argument validation and
 * submission only, and each d_results[k] is valid once the stream passes
 * this call.
 *
 * Per chunk: CUDEC_OK with bytes_written, and exactly bytes_written bytes
 * stored at d_dst_ptrs[k] (nothing beyond them is touched); or
 * CUDEC_ERR_CORRUPT_INPUT / CUDEC_ERR_OUTPUT_TOO_SMALL with bytes_written 0
 * and unspecified destination contents.
 *
 * Returns CUDEC_ERR_INVALID_ARGUMENT, without any device-runtime call and
 * leaving the thread's pending runtime error in place, for a NULL array, a
 * misaligned d_results, chunk_count == 0, or chunk_count above the launch
 * limit. A call that passes validation first consumes the pending runtime
 * error, so CUDEC_ERR_CUDA reports this submission's own failure. */
cudec_status cudec_lz4_decompress_batch(const void* const* d_src_ptrs,
                                        const size_t* d_src_sizes,
                                        void* const* d_dst_ptrs,
                                        const size_t* d_dst_capacities,
                                        size_t chunk_count,
                                        cudec_chunk_result* d_results,
                                        cudec_stream_t stream);

/* Decode one LZ4 frame, host memory to host memory, synchronously, through
 * the batch decoder. Block-independent frames only; header, block and
 * content checksums and the content size are verified when present.
 * CUDEC_ERR_UNSUPPORTED for linked blocks or a dictionary id,
 * CUDEC_ERR_CORRUPT_INPUT for a malformed frame or a failed check,
 * CUDEC_ERR_OUTPUT_TOO_SMALL when dst_capacity is short, CUDEC_ERR_CUDA on a
 * device or host resource failure, CUDEC_ERR_INVALID_ARGUMENT for a NULL
 * frame or bytes_written, or a NULL dst with a non-zero capacity. On any
 * error *bytes_written is 0. */
cudec_status cudec_lz4f_decompress(const void* frame, size_t frame_size,
                                   void* dst, size_t dst_capacity,
                                   size_t* bytes_written);

/* Where the streaming decoder's destinations live. */
typedef enum cudec_mem_space {
    CUDEC_MEM_HOST = 0,  /* host pointers: output is copied back */
    CUDEC_MEM_DEVICE = 1 /* device pointers: output stays on the device */
} cudec_mem_space;

/* A reusable streaming-decode context: it owns a stream and grow-only
 * pinned and device staging. One per thread. */
typedef struct cudec_stream_ctx cudec_stream_ctx;

/* Creates a context (stream and event only; staging is allocated on first
 * use). CUDEC_ERR_INVALID_ARGUMENT for a NULL out_ctx, CUDEC_ERR_CUDA on a
 * resource failure; *out_ctx is NULL on any error. */
cudec_status cudec_stream_ctx_create(cudec_stream_ctx** out_ctx);

/* Synchronous batch LZ4 decode from host-resident compressed chunks.
 * All arrays are host arrays of chunk_count entries; destinations are in
 * dst_space. On return every h_results[k] is valid. A successful chunk
 * leaves exactly bytes_written bytes at dst_ptrs[k] and nothing beyond.
 * Output is identical for a fresh or reused context.
 *
 * CUDEC_ERR_INVALID_ARGUMENT (nothing decoded, context not poisoned) for a
 * NULL ctx or array, a NULL h_src_ptrs[k] with a non-zero size, a NULL
 * dst_ptrs[k] with a non-zero capacity, an unknown dst_space,
 * chunk_count == 0, or a count above the launch limit. Otherwise the
 * aggregate is CUDEC_OK iff every chunk is OK; else CUDEC_ERR_CUDA for a
 * resource or device failure, else the first failing chunk's status in
 * index order. On a non-OK aggregate every h_results[k] holds a defined
 * non-OK status with bytes_written 0 (CUDEC_ERR_CUDA for a chunk the call
 * did not produce). A device fault poisons the context: every later decode
 * returns CUDEC_ERR_CUDA without touching the device. */
cudec_status cudec_lz4_decompress_stream_ctx(
    cudec_stream_ctx* ctx, const void* const* h_src_ptrs,
    const size_t* h_src_sizes, void* const* dst_ptrs, const size_t* dst_caps,
    size_t chunk_count, cudec_mem_space dst_space,
    cudec_chunk_result* h_results);

/* Frees a context and everything it owns; NULL-safe, valid when poisoned. */
void cudec_stream_ctx_destroy(cudec_stream_ctx* ctx);

/* create + one cudec_lz4_decompress_stream_ctx + destroy. */
cudec_status cudec_lz4_decompress_stream(const void* const* h_src_ptrs,
                                         const size_t* h_src_sizes,
                                         void* const* dst_ptrs,
                                         const size_t* dst_caps,
                                         size_t chunk_count,
                                         cudec_mem_space dst_space,
                                         cudec_chunk_result* h_results);

#ifdef __cplusplus
}
#endif

#endif /* CUDEC_H */
