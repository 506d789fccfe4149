/* Driver over iderex/cudec's own ladder headers, compiled from the
 * reference tree where it lies (see oracle/Makefile, target `ref`); the
 * output goes to oracle/_ref/ and is never committed. TEST INFRASTRUCTURE
 * ONLY: it is a second, independent checker for oracle/cudec_oracle.c.
 *
 * The headers are the reference's src/lz4_block.h and src/snappy_block.h,
 * included unmodified. What is written here is only the execution loop,
 * following the reference's host twins (tests/lz4_twin.h:41-77,
 * tests/snappy_twin.h:140-214) and the kernel's fuel / result rules
 * (src/lz4_decode.cuh:90-171). */
#include "cudec.h"
#include "lz4_block.h"
#include "snappy_block.h"

#include <cstdint>

extern "C" {

int32_t ref_lz4_decode(const uint8_t* src, uint64_t src_size, uint8_t* dst,
                       uint64_t cap, uint64_t* written) {
    cudec_detail::Lz4Parser parser{src, src_size, cap};
    cudec_detail::Lz4Sequence seq;
    bool done = false;
    cudec_status status = CUDEC_OK;
    uint64_t fuel = src_size == UINT64_MAX ? UINT64_MAX : src_size + 1;
    for (;;) {
        status = parser.Next(&seq, &done);
        if (status != CUDEC_OK) {
            break;
        }
        for (uint64_t i = 0; i < seq.literals_len; i++) {
            dst[seq.literals_dst + i] = src[seq.literals_src + i];
        }
        for (uint64_t i = 0; i < seq.match_len; i++) {
            dst[seq.match_dst + i] = dst[seq.match_src + i];
        }
        if (done || fuel-- == 0) {
            break;
        }
    }
    if (status == CUDEC_OK && !done) {
        status = CUDEC_ERR_CORRUPT_INPUT;
    }
    *written = status == CUDEC_OK ? parser.dst_pos : 0;
    return status;
}

int32_t ref_lz4_parse(const uint8_t* src, uint64_t src_size, uint64_t cap,
                      uint64_t* dst_pos) {
    cudec_detail::Lz4Parser parser{src, src_size, cap};
    cudec_detail::Lz4Sequence seq;
    bool done = false;
    cudec_status status = CUDEC_OK;
    uint64_t fuel = src_size == UINT64_MAX ? UINT64_MAX : src_size + 1;
    for (;;) {
        status = parser.Next(&seq, &done);
        if (status != CUDEC_OK || done || fuel-- == 0) {
            break;
        }
    }
    if (status == CUDEC_OK && !done) {
        status = CUDEC_ERR_CORRUPT_INPUT;
    }
    *dst_pos = parser.dst_pos;
    return status;
}

int32_t ref_snappy_decode(const uint8_t* src, uint64_t src_size, uint8_t* dst,
                          uint64_t cap, uint64_t* written, int32_t* reject) {
    cudec_detail::SnappyParser parser{src, src_size, cap};
    cudec_detail::SnappyElement element;
    bool done = false;
    cudec_status status = CUDEC_OK;
    uint64_t fuel = src_size == UINT64_MAX ? UINT64_MAX : src_size + 1;
    for (;;) {
        status = parser.Next(&element, &done);
        if (status != CUDEC_OK) {
            break;
        }
        if (element.is_copy) {
            for (uint64_t i = 0; i < element.length; i++) {
                dst[element.to + i] = dst[element.from + i];
            }
        } else {
            for (uint64_t i = 0; i < element.length; i++) {
                dst[element.to + i] = src[element.from + i];
            }
        }
        if (done || fuel-- == 0) {
            break;
        }
    }
    if (status == CUDEC_OK && !done) {
        status = CUDEC_ERR_CORRUPT_INPUT;
    }
    if (reject) {
        *reject = status == CUDEC_OK ? 0 : static_cast<int32_t>(parser.reject);
    }
    *written = status == CUDEC_OK ? parser.dst_pos : 0;
    return status;
}

int32_t ref_snappy_declared(const uint8_t* src, uint64_t src_size,
                            uint64_t* declared, uint64_t* preamble_size,
                            int32_t* reject) {
    cudec_detail::SnappyReject r = cudec_detail::kSnappyRejectNone;
    const cudec_status st = cudec_detail::SnappyDeclaredLength(
        src, src_size, declared, preamble_size, &r);
    if (reject) {
        *reject = static_cast<int32_t>(r);
    }
    return st;
}

}  // extern "C"
