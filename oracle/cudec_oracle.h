/* CPU oracle for the batch LZ4 / Snappy block decoders.
 *
 * TEST INFRASTRUCTURE ONLY. Nothing in the product links, loads or calls
 * this code: it is the checker that the HIP decoders are held to (tests,
 * the smoke check, and the bench's CPU baseline leg).
 *
 * It is a sequential plain-C restatement of iderex/cudec's validation
 * ladders (src/lz4_block.h, src/snappy_block.h), executed the way the
 * reference's host twins execute them (tests/lz4_twin.h,
 * tests/snappy_twin.h): sequence by sequence, with a chasing byte copy for
 * overlapping matches, and with the kernel's fuel and result semantics
 * (src/lz4_decode.cuh:90-171).
 *
 * Status values are cudec_status values (0 OK, 2 CORRUPT_INPUT,
 * 3 OUTPUT_TOO_SMALL). */
#ifndef CUDEC_ORACLE_H
#define CUDEC_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* The Snappy ladder's reject rungs, in the reference's enumeration order
 * (src/snappy_block.h:118-134). */
enum oracle_snappy_reject {
    ORACLE_SNAPPY_REJECT_NONE = 0,
    ORACLE_SNAPPY_REJECT_PREAMBLE_BYTE_MISSING,
    ORACLE_SNAPPY_REJECT_PREAMBLE_FINAL_GROUP,
    ORACLE_SNAPPY_REJECT_PREAMBLE_UNREACHABLE,
    ORACLE_SNAPPY_REJECT_DECLARED_OVER_CAPACITY,
    ORACLE_SNAPPY_REJECT_PAST_TERMINAL,
    ORACLE_SNAPPY_REJECT_UNDER_PRODUCTION,
    ORACLE_SNAPPY_REJECT_LITERAL_HEADER_TRUNCATED,
    ORACLE_SNAPPY_REJECT_LITERAL_PAYLOAD_TRUNCATED,
    ORACLE_SNAPPY_REJECT_LITERAL_OVER_PRODUCTION,
    ORACLE_SNAPPY_REJECT_COPY_HEADER_TRUNCATED,
    ORACLE_SNAPPY_REJECT_COPY_OFFSET_ZERO,
    ORACLE_SNAPPY_REJECT_COPY_OFFSET_BEFORE_START,
    ORACLE_SNAPPY_REJECT_COPY_OVER_PRODUCTION,
    ORACLE_SNAPPY_REJECT_COUNT
};

/* One LZ4 block into dst[0, cap). Returns the status; *written is the
 * decoded size on OK and 0 otherwise. dst may be NULL only when cap == 0.
 * On a reject dst holds whatever the ladder had authorised (unspecified). */
int32_t oracle_lz4_decode(const uint8_t* src, uint64_t src_size, uint8_t* dst,
                          uint64_t cap, uint64_t* written);

/* The ladder alone, no copies: any capacity (SIZE_MAX included). Reports the
 * final dst_pos, the number of sequences, and the number of calls that
 * returned OK without advancing src_pos (must be 0: liveness). */
int32_t oracle_lz4_parse(const uint8_t* src, uint64_t src_size, uint64_t cap,
                         uint64_t* dst_pos, uint64_t* sequences,
                         uint64_t* stalls);

/* The Snappy preamble alone (src/snappy_block.h:190-223). */
int32_t oracle_snappy_declared(const uint8_t* src, uint64_t src_size,
                               uint64_t* declared, uint64_t* preamble_size,
                               int32_t* reject);

/* One Snappy block into dst[0, cap), the batch contract: the declared
 * length is checked against cap (OUTPUT_TOO_SMALL), success needs exact
 * consumption and exact production. *reject names the refusing rung. */
int32_t oracle_snappy_decode(const uint8_t* src, uint64_t src_size,
                             uint8_t* dst, uint64_t cap, uint64_t* written,
                             int32_t* reject);

/* The Snappy ladder alone, no copies, with the liveness count. */
int32_t oracle_snappy_parse(const uint8_t* src, uint64_t src_size,
                            uint64_t cap, uint64_t* dst_pos,
                            uint64_t* elements, uint64_t* stalls,
                            int32_t* reject);

/* Decodes n chunks on `threads` host threads (static contiguous partition)
 * and returns the wall-clock seconds of the decode region alone (threads
 * are started and parked before the clock starts). format 0 = LZ4,
 * 1 = Snappy. status / written receive the per-chunk results. */
double oracle_decode_batch_timed(int32_t format, const uint8_t* const* srcs,
                                 const uint64_t* src_sizes,
                                 uint8_t* const* dsts, const uint64_t* caps,
                                 uint64_t n, int32_t threads, int32_t* status,
                                 uint64_t* written);

#ifdef __cplusplus
}
#endif

#endif /* CUDEC_ORACLE_H */
