/* Device-side LZ4 and Snappy block ladders for gfx950.
 *
 * These are the fail-closed validation ladders every kernel form runs: all
 * stream-derived values are checked before first use, lengths accumulate in
 * 64 bits with the remaining capacity re-checked, and success is reported
 * only at exact consumption. The rules and their order follow iderex/cudec's
 * src/lz4_block.h:46-258 and src/snappy_block.h:58-435; status values are
 * cudec_status values.
 *
 * The cursor state is wave-uniform: every lane of the wave (or sub-group)
 * that owns a chunk holds the same values, and the kernels pin that with
 * readfirstlane so the parse runs on the scalar unit. */
#ifndef CUDEC_AMD_LADDER_H
#define CUDEC_AMD_LADDER_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "common.h"
#include "cudec.h"

namespace cudec_amd {

constexpr int kOk = CUDEC_OK;
constexpr int kCorrupt = CUDEC_ERR_CORRUPT_INPUT;
constexpr int kTooSmall = CUDEC_ERR_OUTPUT_TOO_SMALL;

/* LZ4 constants (src/lz4_block.h:46-60). */
constexpr uint64_t kLz4SlackDst = 12;   /* MFLIMIT */
constexpr uint64_t kLz4SlackSrc = 8;    /* offset + token + LASTLITERALS */
constexpr uint64_t kLz4LastLiterals = 5;
constexpr uint64_t kLz4RunMask = 15;
constexpr uint64_t kLz4MinMatch = 4;
constexpr uint64_t kLz4LitMargin = 15;  /* liblz4 ilimit, literal lengths */
constexpr uint64_t kLz4MatchMargin = 4; /* liblz4 ilimit, match lengths */

struct Lz4Cursor {
    gcu8* src;
    uint64_t src_size;
    uint64_t cap;
    uint64_t sp;
    uint64_t dp;
};

/* One executable step: copy lit_len bytes src[lit_src..] -> dst[lit_dst..],
 * then match_len bytes dst[match_dst - offset + (i mod offset)] ->
 * dst[match_dst + i]. match_len == 0 on the terminal step. */
struct Lz4Step {
    uint64_t lit_src;
    uint64_t lit_dst;
    uint64_t lit_len;
    uint64_t match_dst;
    uint64_t match_len;
    uint32_t offset;
};

/* Where a ladder reads its bytes. GlobalBytes: the chunk's source in
 * global memory. StagedBytes: a copy of src[base, base + avail) in LDS
 * (a window's stage), falling back to global memory past it; the ladder
 * only reads at or after base. Both are read with wave-uniform positions. */
struct GlobalBytes {
    gcu8* src;
    __device__ __forceinline__ uint32_t byte(uint64_t p) const { return src[p]; }
    __device__ __forceinline__ bool local(uint64_t, uint32_t) const { return false; }
    __device__ __forceinline__ uint32_t word(uint64_t) const { return 0; }
};
struct StagedBytes {
    gcu8* src;
    const uint8_t* cw;
    uint64_t base;
    uint32_t avail;
    __device__ __forceinline__ uint32_t byte(uint64_t p) const {
        return p - base < avail ? (uint32_t)cw[p - base] : (uint32_t)src[p];
    }
    /* the n bytes at p are all in the stage */
    __device__ __forceinline__ bool local(uint64_t p, uint32_t n) const {
        return p - base < avail && avail - (p - base) >= n;
    }
    __device__ __forceinline__ uint32_t word(uint64_t p) const {
        uint32_t v;
        __builtin_memcpy(&v, cw + (p - base), 4);
        return v;
    }
};

/* 255-continued length extension (src/lz4_block.h:95-128). */
__device__ __forceinline__ int lz4_accumulate(Lz4Cursor& c, uint64_t& length,
                                              uint64_t margin,
                                              bool initial_check) {
    if (initial_check && c.sp + margin >= c.src_size) {
        return kCorrupt;
    }
    uint64_t fuel = c.src_size - c.sp + 1;
    uint32_t byte;
    do {
        if (fuel-- == 0) {
            return kCorrupt;
        }
        if (c.sp >= c.src_size) {
            return kCorrupt;
        }
        if (length > c.cap - c.dp) {
            return kTooSmall;
        }
        byte = c.src[c.sp++];
        length += byte;
        if (c.sp + margin > c.src_size) {
            return kCorrupt;
        }
    } while (byte == 255);
    return kOk;
}

/* lz4_accumulate with the whole wave reading continuation bytes, 256 a
 * round: lane i takes the four bytes at sp + 4 i (one dword load where all
 * four lie in the source). Each lane evaluates the serial loop's checks for
 * its bytes as that loop would meet them there (every earlier byte of the
 * run is 255, so the length before byte k is length + 255 k), and the
 * first byte in stream order at which the loop would leave decides, with
 * the status the loop leaves with. Results equal lz4_accumulate's. Its fuel
 * check never decides: it could first fire one byte after the one where sp
 * reaches src_size, and that byte already returns. Called with the whole
 * wave active and the cursor uniform. */
template <class Bytes>
__device__ __forceinline__ int lz4_accumulate_wave(Lz4Cursor& c, const Bytes& in, uint64_t& length,
                                                   uint64_t margin, bool initial_check) {
    if (initial_check && c.sp + margin >= c.src_size) {
        return kCorrupt;
    }
    const uint32_t lane = __lane_id();
    const uint64_t room = c.cap - c.dp;
    constexpr int kNone = -1;
    for (;;) {
        const uint64_t p0 = c.sp + 4u * lane;
        uint32_t bytes = 0;
        if (p0 + 4 <= c.src_size && in.local(p0, 4)) {
            bytes = in.word(p0);
        } else if (p0 + 4 <= c.src_size) {
            __builtin_memcpy(&bytes, (const void*)(c.src + p0), 4);
        } else {
            for (uint32_t u = 0; u < 4; u++) {
                if (p0 + u < c.src_size) {
                    bytes |= (uint32_t)c.src[p0 + u] << (8 * u);
                }
            }
        }
        int st = kNone;
        uint32_t at = 0;
#pragma unroll
        for (uint32_t u = 0; u < 4; u++) {
            const uint64_t p = p0 + u;
            const uint32_t byte = (bytes >> (8 * u)) & 0xFFu;
            if (st == kNone) {
                at = u;
                if (p >= c.src_size) {
                    st = kCorrupt;
                } else if (length + 255u * (p - c.sp) > room) {
                    st = kTooSmall;
                } else if (p + 1 + margin > c.src_size) {
                    st = kCorrupt;
                } else if (byte != 255u) {
                    st = kOk;
                }
            }
        }
        const uint64_t hit = __ballot(st != kNone);
        if (hit != 0) {
            const uint32_t i = (uint32_t)__builtin_ctzll(hit);
            const int s = (int)uniform32(__builtin_amdgcn_readlane((uint32_t)st, i));
            if (s == kOk) {
                const uint32_t u = uniform32(__builtin_amdgcn_readlane(at, i));
                const uint32_t word = uniform32(__builtin_amdgcn_readlane(bytes, i));
                const uint64_t k = 4u * (uint64_t)i + u;
                length += 255u * k + ((word >> (8 * u)) & 0xFFu);
                c.sp += k + 1;
            }
            return s;
        }
        length += 255u * 256u;
        c.sp += 256;
    }
}

/* One sequence (src/lz4_block.h:139-258), its header bytes read from `in`.
 * kWaveScan: length extensions are read by the whole wave
 * (lz4_accumulate_wave), for callers that run the ladder with every lane
 * active; otherwise byte by byte from global memory. */
template <bool kWaveScan = false, class Bytes = GlobalBytes>
__device__ __forceinline__ int lz4_next(Lz4Cursor& c, Lz4Step& q,
                                        bool& done, const Bytes& in) {
    done = false;
    if (c.sp >= c.src_size) {
        return kCorrupt;
    }
    const uint32_t token = in.byte(c.sp++);
    uint64_t lit = token >> 4;
    if (lit == kLz4RunMask) {
        const int st = kWaveScan ? lz4_accumulate_wave(c, in, lit, kLz4LitMargin, true)
                                 : lz4_accumulate(c, lit, kLz4LitMargin, true);
        if (st != kOk) {
            return st;
        }
    }
    if (lit > c.src_size - c.sp) {
        return kCorrupt;
    }
    if (lit > c.cap - c.dp) {
        return kTooSmall;
    }
    q.lit_src = c.sp;
    q.lit_dst = c.dp;
    q.lit_len = lit;
    const uint64_t dst_slack = c.cap - c.dp - lit;
    const uint64_t src_slack = c.src_size - c.sp - lit;
    if (dst_slack < kLz4SlackDst || src_slack < kLz4SlackSrc) {
        if (c.sp + lit != c.src_size) {
            return kCorrupt;
        }
        if (c.cap == 0 && token != 0) {
            return kCorrupt;
        }
        c.sp += lit;
        c.dp += lit;
        q.match_dst = c.dp;
        q.match_len = 0;
        q.offset = 1;
        done = true;
        return kOk;
    }
    c.sp += lit;
    c.dp += lit;
    const uint32_t offset = in.byte(c.sp) | (in.byte(c.sp + 1) << 8);
    c.sp += 2;
    if (offset == 0) {
        return kCorrupt;
    }
    if (offset > c.dp) {
        return kCorrupt;
    }
    uint64_t match = (token & kLz4RunMask) + kLz4MinMatch;
    if ((token & kLz4RunMask) == kLz4RunMask) {
        const int st = kWaveScan ? lz4_accumulate_wave(c, in, match, kLz4MatchMargin, false)
                                 : lz4_accumulate(c, match, kLz4MatchMargin, false);
        if (st != kOk) {
            return st;
        }
    }
    if (match > c.cap - c.dp) {
        return kTooSmall;
    }
    if (match > c.cap - c.dp - kLz4LastLiterals) {
        return kCorrupt;
    }
    q.match_dst = c.dp;
    q.match_len = match;
    q.offset = offset;
    c.dp += match;
    return kOk;
}

__device__ __forceinline__ int lz4_next(Lz4Cursor& c, Lz4Step& q, bool& done) {
    return lz4_next<false>(c, q, done, GlobalBytes{c.src});
}

/* Snappy (src/snappy_block.h:58-435). */
struct SnappyCursor {
    gcu8* src;
    uint64_t src_size;
    uint64_t cap;
    uint64_t sp;
    uint64_t dp;
    uint64_t declared;
};

/* One element: a literal (is_copy false: src[from..] -> dst[to..]) or a
 * copy (dst[to - offset + (i mod offset)] -> dst[to + i]). */
struct SnappyStep {
    uint64_t from;
    uint64_t to;
    uint64_t length;
    uint64_t offset;
    bool is_copy;
};

/* Preamble and the declared length against the capacity
 * (src/snappy_block.h:190-263). */
__device__ __forceinline__ int snappy_begin(SnappyCursor& c) {
    uint64_t result = 0;
    for (uint32_t g = 0; g < 5; g++) {
        if (g >= c.src_size) {
            return kCorrupt;
        }
        const uint32_t b = c.src[g];
        if (g == 4 && b >= 16) {
            return kCorrupt;
        }
        result |= (uint64_t)(b & 0x7F) << (7 * g);
        if ((b & 0x80) == 0) {
            if (result > c.cap) {
                return kTooSmall;
            }
            c.declared = result;
            c.sp = g + 1;
            c.dp = 0;
            return kOk;
        }
    }
    return kCorrupt;
}

/* One element after a successful begin (src/snappy_block.h:288-435). */
__device__ __forceinline__ int snappy_next(SnappyCursor& c, SnappyStep& e,
                                           bool& done) {
    done = false;
    if (c.sp >= c.src_size) {
        if (c.dp != c.declared) {
            return kCorrupt;
        }
        e.from = 0;
        e.to = c.dp;
        e.length = 0;
        e.offset = 1;
        e.is_copy = false;
        done = true;
        return kOk;
    }
    const uint32_t tag = c.src[c.sp];
    const uint32_t kind = tag & 3u;
    if (kind == 0) {
        uint64_t length = tag >> 2;
        uint64_t header = 1;
        if (length > 59) {
            const uint64_t extra = length - 59;
            header += extra;
            if (header > c.src_size - c.sp) {
                return kCorrupt;
            }
            length = 0;
            for (uint64_t i = 0; i < extra; i++) {
                length |= (uint64_t)c.src[c.sp + 1 + i] << (8 * i);
            }
        }
        length += 1;
        c.sp += header;
        if (length > c.src_size - c.sp) {
            return kCorrupt;
        }
        if (length > c.declared - c.dp) {
            return kCorrupt;
        }
        e.from = c.sp;
        e.to = c.dp;
        e.length = length;
        e.offset = 1;
        e.is_copy = false;
        c.sp += length;
        c.dp += length;
        return kOk;
    }
    const uint64_t offset_bytes = kind == 1 ? 1 : (kind == 2 ? 2 : 4);
    if (1 + offset_bytes > c.src_size - c.sp) {
        return kCorrupt;
    }
    uint64_t offset = 0;
    for (uint64_t i = 0; i < offset_bytes; i++) {
        offset |= (uint64_t)c.src[c.sp + 1 + i] << (8 * i);
    }
    uint64_t length;
    if (kind == 1) {
        offset |= (uint64_t)(tag >> 5) << 8;
        length = 4 + ((tag >> 2) & 7u);
    } else {
        length = (uint64_t)(tag >> 2) + 1;
    }
    c.sp += 1 + offset_bytes;
    if (offset == 0) {
        return kCorrupt;
    }
    if (offset > c.dp) {
        return kCorrupt;
    }
    if (length > c.declared - c.dp) {
        return kCorrupt;
    }
    e.from = c.dp - offset;
    e.to = c.dp;
    e.length = length;
    e.offset = offset;
    e.is_copy = true;
    c.dp += length;
    return kOk;
}

}  // namespace cudec_amd

#endif /* CUDEC_AMD_LADDER_H */
