/* The tuned launch form of the LZ4 and Snappy batch decoders for gfx950.
 *
 * Work distribution: four small kernels bin the chunks by decode cost
 * (capacity + (capacity - compressed size), the output plus what matches
 * add to it; four bins per octave) into a device queue. A per-device,
 * per-format decode plan names, for each cost bin, the lanes that decode
 * one of its chunks: 128 (both waves of a block, as a software pipeline)
 * or 64 (one wave). The queue holds the two-wave chunks first, then the
 * one-wave ones, each class costliest first. A persistent grid of two-wave
 * blocks pulls chunks from it; once the queue reaches the one-wave class,
 * a block splits into two independent one-wave decoders.
 *
 * Per chunk, the stream is decoded in windows of candidate positions
 * staged in LDS:
 *   1. stage     -- compressed bytes from the window start are loaded as
 *                   aligned 16-byte blocks (a block holding any byte of the
 *                   chunk never leaves the chunk's pages);
 *   2. speculate -- every lane parses 16 (two-wave form) or 8 (one-wave
 *                   form) candidate positions as if a sequence
 *                   (LZ4) or element (Snappy) started there, and records
 *                   where the next one would start;
 *   3. walk      -- a scalar walk follows the true chain from position 0
 *                   through those links (v_readlane): up to 64 members, one
 *                   per lane;
 *   4. resolve   -- a wavefront prefix sum places each member's output and
 *                   the exact validation ladder (src/lz4_block.h:95-258,
 *                   src/snappy_block.h:288-435) runs per lane; the first
 *                   member in stream order that rejects, ends the stream,
 *                   or would overflow the group ends the group, so a chunk's
 *                   status is the sequential ladder's;
 *   5. execute   -- output goes through an LDS ring: literals by member
 *                   lanes; matches whose sources are final by member lanes
 *                   (batched 16-byte history loads); the rest in stream
 *                   order with the wave fanned out over the bytes;
 *   6. flush     -- completed ring bytes leave as aligned 16-byte stores;
 *                   only a chunk's first and last partial blocks use byte
 *                   stores, so exactly bytes_written bytes are stored.
 * Steps 1-4 read only the source and decide every status; steps 5-6 only
 * write. In the two-wave form one wave runs 1-4 on window i + 1 while the
 * other runs 5-6 on window i. A sequence the window cannot hold (a length
 * with two or more extension bytes, a Snappy literal over 256 bytes) is
 * decided by the exact scalar ladder and copied wave-wide, global to global.
 */
#include <cstdlib>
#include <cstring>
#include <mutex>

#include "common.h"
#include "cudec_mi355x.h"
#include "ladder.h"

namespace cudec_amd {
namespace {

/* Candidates per lane: a window spans kC * 64 candidate positions. The
 * LZ4 two-wave form parses wider windows than the one-wave form: its
 * parser runs beside the copier, and a wider window fills more lanes of
 * every lane-parallel step, while a lone wave pays the wider speculation
 * on its own critical path. Snappy's two-wave windows are 448 candidates:
 * a window holds at most 64 members, and Snappy elements (2 to 5 bytes of
 * the source, mostly) fill those within a few hundred bytes, so wider
 * windows only add speculation to the busier parser wave (768: 35.2 ms,
 * 576: 32.1, 448: 30.7, 384: 31.9 on the uneven batch); its two-wave path
 * fits 96 VGPRs, so Snappy launches run five waves per SIMD throughout. */
constexpr uint32_t kCandSplit = 8;
constexpr uint32_t kCandPipe = 16;
constexpr uint32_t kCandPipeSnappy = 7;
/* staged bytes a window of kC candidates per lane reads: the window plus
 * the longest candidate that starts in it (an LZ4 sequence with one
 * extension byte per length reads 275 bytes, a Snappy literal window 277),
 * in whole 16-byte blocks. (A 512-byte margin staged 1.5 KiB a window; this
 * one takes the one-wave form 2.8% faster.) */
constexpr uint32_t stage_bytes(uint32_t kc) { return kc * 64u + 304u; }
constexpr uint32_t kStageBuf = stage_bytes(kCandPipe) + 16;
constexpr uint32_t kRing = 4096;          /* per-wave output ring */
constexpr uint32_t kBatchMax = 2048;      /* output bytes of one batch, one-wave form */
constexpr uint32_t kStop = 0xFFFFu;       /* link of an unresolvable candidate */
constexpr uint32_t kLaneLit = 64;         /* literal runs copied by their lane */
constexpr uint32_t kLaneMatch = 64;       /* matches copied by their lane (<= 64) */
constexpr uint32_t kOctaveBits = 2;       /* cost bins per octave: 4 */
constexpr int kBins = 64 << kOctaveBits;
static_assert(kBins == kCostBins, "the plan API names every bin");
/* Default plan: chunks costing at least 512 KiB (capacity + capacity -
 * compressed) run on two waves. */
constexpr uint64_t kDefaultPipeMinCost = 512u << 10;
/* Queue keys: a chunk's class (two-wave first), then its cost bin. */
constexpr int kKeys = 2 * kBins;

static_assert(kBatchMax + 32 <= kRing, "ring must hold a batch and a margin");
static_assert((kRing & (kRing - 1)) == 0, "rings are powers of two");
static_assert(kLaneMatch <= 64, "a lane match copies at most four 16-byte blocks");
static_assert(kCandSplit * 64 + 2 + 269 + 3 <= stage_bytes(kCandSplit) &&
                  kCandPipe * 64 + 2 + 269 + 3 <= stage_bytes(kCandPipe),
              "a window's last candidate fits the stage");

/* A wave's output ring (kR bytes, a power of two) and a stage buffer. The
 * one-wave decoders keep a 4 KiB ring each and the link table of their
 * window (walk_doubling); the two-wave form's copier an 8 KiB ring in the
 * same LDS (the two never run at once in a block). */
template <uint32_t kR>
struct alignas(16) RingLds {
    static constexpr uint32_t kSize = kR;
    uint8_t ring[kR];
    uint8_t stage[kStageBuf];
    uint8_t sink[kWave]; /* stores of lanes with nothing to store */
};
struct alignas(16) WaveLds : RingLds<kRing> {
    uint16_t jump[kCandSplit * kWave];
};
constexpr uint32_t kPipeRing = 8192;
using PipeRing = RingLds<kPipeRing>;
/* Output bytes of one batch in the two-wave form. Windows of token data
 * produce under 2 KiB; a window of long short-offset matches (64 members
 * of about 90 bytes each) fills up to 6 KiB, which its 8 KiB ring holds
 * with 2 KiB of history. */
constexpr uint32_t kBatchMaxPipe = 6144;
static_assert(kBatchMaxPipe + 32 <= kPipeRing, "the two-wave ring holds a batch and a margin");
/* a batch's limit: the one-wave form walks by pointer doubling */
constexpr uint32_t batch_max(bool one_wave) { return one_wave ? kBatchMax : kBatchMaxPipe; }

/* The decode queue hands out positions of the cost-sorted chunk order in
 * up to kSubQueues interleaved sub-queues (of k: sub-queue s holds
 * positions s, s + k, ...), each counter on its own 4 KiB page: the
 * counters' atomics are served in parallel, where one counter for the
 * whole grid serialises every chunk's pop (about 12 ns each, the pace of a
 * batch of many small chunks). A batch of few chunks per wave keeps one
 * queue: its exact costliest-first order matters more there. */
constexpr uint32_t kSubQueues = 8;
constexpr uint32_t kSubStride = 1024; /* uint32 words: 4 KiB apart */
constexpr uint64_t kSubQueueMinPerWave = 16; /* chunks per wave for kSubQueues */

/* A wave's place in the decode queue: the sub-queue it draws from and how
 * many sub-queues it has not yet found drained. */
struct QueueCursor {
    uint32_t sub;
    uint32_t left;
};

/* Scratch of one tuned launch (stream-ordered allocation). */
struct Queue {
    uint32_t next[kSubQueues * kSubStride];
    uint32_t bin_count[kKeys];
    uint32_t bin_base[kKeys];
    uint32_t bin_cursor[kKeys];
};

/* A decode plan: bit b (of word b / 64) set when cost bin b decodes on
 * two pipelined waves, clear for one wave. Passed to the kernels by value. */
struct Plan {
    uint64_t two_wave[kBins / 64];
};

__host__ __device__ inline bool plan_two_wave(const Plan& p, uint32_t bin) {
    return ((p.two_wave[bin >> 6] >> (bin & 63)) & 1u) != 0;
}

template <uint32_t kR = kRing>
__device__ __forceinline__ uint32_t ring_index(uintptr_t dst_addr, uint64_t pos) {
    return (uint32_t)(dst_addr + pos) & (kR - 1);
}

/* ------------------------------------------------------------ staging */

/* Loads compressed bytes [sp, sp + stage_bytes(kC)) of the chunk into stage and
 * returns the offset of byte sp inside it. Only aligned 16-byte blocks that
 * start before the chunk end are read (each holds at least one byte of the
 * chunk, so none leaves the chunk's pages); later blocks are zero-filled.
 * Bytes past src_size are never trusted: every member is validated against
 * src_size before anything it read is used. */
template <uint32_t kC>
__device__ __forceinline__ uint32_t stage_into(uint8_t* stage, gcu8* src,
                                               uint64_t src_size, uint64_t sp,
                                               uint32_t lane) {
    const uintptr_t a = (uintptr_t)(src + sp);
    const uintptr_t base = a & ~(uintptr_t)15;
    const uintptr_t lo = (uintptr_t)src;
    const uintptr_t end = lo + src_size;
    for (uint32_t blk = lane; blk < stage_bytes(kC) / 16 + 1; blk += kWave) {
        const uintptr_t g = base + 16u * blk;
        uint4 v = make_uint4(0, 0, 0, 0);
        /* only a block that holds at least one byte of [src, src_size) */
        if (g < end && g + 16 > lo) {
            v = gload16((gcu8*)g);
        }
        *reinterpret_cast<uint4*>(&stage[16u * blk]) = v;
    }
    wave_order();
    return (uint32_t)(a - base);
}

/* Loads one dword per lane of the 1 KiB after the window staged at sp
 * (16 bytes apart, every one inside the source), so that the next window's
 * stage finds those bytes in L2. The caller holds the value until the
 * window is parsed, which keeps the loads off its critical path. */
template <uint32_t kC>
__device__ __forceinline__ uint32_t touch_ahead(gcu8* src, uint64_t src_size, uint64_t sp,
                                                uint32_t lane) {
    const uintptr_t base = (uintptr_t)(src + sp) & ~(uintptr_t)15;
    const uintptr_t g = base + 16u * (stage_bytes(kC) / 16 + 1) + 16u * lane;
    const uintptr_t end = (uintptr_t)src + src_size;
    return g + 4 <= end ? *(gcu32*)g : 0u;
}

/* ------------------------------------------------------ candidate parse */

struct Cand {
    uint32_t tok;
    uint32_t lit_len;
    uint32_t lit;    /* stage-relative position of the literals */
    uint32_t offset;
    uint32_t match_len;
    uint32_t next;   /* position after the sequence */
    bool complex;    /* a length needs two or more extension bytes */
};

/* The sequence a token at window position c would encode, assuming at most
 * one extension byte per length. Reads stay below c + 275, inside the
 * stage. (Windows that also held matches with two extension bytes, up to
 * 529 bytes, decoded runs of long short-offset matches 20% faster but token
 * data 2.4% slower: the wider parse cost the kernel registers it spilled.) */
__device__ __forceinline__ Cand parse_candidate(const uint8_t* cw, uint32_t c) {
    /* branch-free: the token and the byte after it, then the three bytes
     * at the offset (offset, match extension); bytes past the sequence are
     * read but never used */
    Cand k;
    k.tok = cw[c];
    const uint32_t e1 = cw[c + 1];
    const uint32_t lh = k.tok >> 4;
    const bool x1 = lh == 15;
    k.lit_len = lh + (x1 ? e1 : 0);
    k.lit = c + 1 + (x1 ? 1 : 0);
    const uint32_t p = k.lit + k.lit_len;
    k.offset = (uint32_t)cw[p] | ((uint32_t)cw[p + 1] << 8);
    const uint32_t e2 = cw[p + 2];
    const uint32_t mh = k.tok & 15;
    const bool x2 = mh == 15;
    k.match_len = mh + 4 + (x2 ? e2 : 0);
    k.next = p + 2 + (x2 ? 1 : 0);
    k.complex = (x1 && e1 == 255) || (x2 && e2 == 255);
    return k;
}

/* ------------------------------------------------------------ wave scan */

__device__ __forceinline__ uint32_t wave_inclusive_sum(uint32_t v, uint32_t lane) {
    /* DPP: an inclusive scan inside each row of 16 lanes (row_shr 1, 2, 4,
     * 8; a lane with no source adds the 0 it keeps), then lane 15 of each
     * row into rows 1 and 3 (row_bcast:15) and lane 31 into rows 2 and 3
     * (row_bcast:31). No LDS traffic. */
    (void)lane;
    int x = (int)v;
    x += __builtin_amdgcn_update_dpp(0, x, 0x111, 0xF, 0xF, false); /* row_shr:1 */
    x += __builtin_amdgcn_update_dpp(0, x, 0x112, 0xF, 0xF, false); /* row_shr:2 */
    x += __builtin_amdgcn_update_dpp(0, x, 0x114, 0xF, 0xF, false); /* row_shr:4 */
    x += __builtin_amdgcn_update_dpp(0, x, 0x118, 0xF, 0xF, false); /* row_shr:8 */
    x += __builtin_amdgcn_update_dpp(0, x, 0x142, 0xA, 0xF, false); /* row_bcast:15 */
    x += __builtin_amdgcn_update_dpp(0, x, 0x143, 0xC, 0xF, false); /* row_bcast:31 */
    return (uint32_t)x;
}

/* ----------------------------------------------------- ring <-> global */

/* Stores ring bytes of positions [from, to) to dst. Full aligned 16-byte
 * blocks go out as one vector store per lane; the partial blocks at either
 * end (only at the chunk's first byte or its last) as single bytes. */
template <class W>
__device__ __forceinline__ void flush_ring(const W& w, gu8* dst,
                                           uint64_t from, uint64_t to,
                                           uint32_t lane) {
    if (to <= from) {
        return;
    }
    const uintptr_t da = (uintptr_t)dst;
    const uintptr_t a0 = da + from;
    const uintptr_t a1 = da + to;
    uintptr_t b0 = (a0 + 15) & ~(uintptr_t)15;
    uintptr_t b1 = a1 & ~(uintptr_t)15;
    if (b0 > a1) {
        b0 = a1;
    }
    if (b1 < b0) {
        b1 = b0;
    }
    /* head bytes [a0, b0) and tail bytes [b1, a1): fewer than 16 each */
    const uint32_t head = (uint32_t)(b0 - a0);
    const uint32_t tail = (uint32_t)(a1 - b1);
    if (lane < head) {
        const uintptr_t g = a0 + lane;
        *(gu8*)g = w.ring[g & (W::kSize - 1)];
    } else if (lane >= 16 && lane < 16 + tail) {
        const uintptr_t g = b1 + (lane - 16);
        *(gu8*)g = w.ring[g & (W::kSize - 1)];
    }
    const uint32_t blocks = (uint32_t)((b1 - b0) / 16);
    for (uint32_t b = lane; b < blocks; b += kWave) {
        const uintptr_t g = b0 + 16u * b;
        const uint4 v = *reinterpret_cast<const uint4*>(&w.ring[g & (W::kSize - 1)]);
        gstore16((gu8*)g, v);
    }
    wave_order();
}

/* One output byte of history: from the ring when its position is still
 * there, else from global memory (already flushed by this wave). The
 * global read takes the aligned 4-byte word holding the byte: a different
 * load width keeps the compiler from merging the two branches into one
 * flat load, and an aligned word never leaves the page of the byte. */
template <class W>
__device__ __forceinline__ uint8_t history_byte(const W& w, gcu8* dst,
                                                uint64_t y, uint64_t ring_lo) {
    if (y >= ring_lo) {
        return w.ring[ring_index<W::kSize>((uintptr_t)dst, y)];
    }
    const uintptr_t a = (uintptr_t)(dst + y);
    const uint32_t word = *(gcu32*)(a & ~(uintptr_t)3);
    return (uint8_t)(word >> ((a & 3) * 8));
}

/* Byte i (0..15) of a 16-byte register block. */
__device__ __forceinline__ uint32_t block_byte(const uint4& v, uint32_t i) {
    const uint32_t x = (i >> 2) == 0 ? v.x : (i >> 2) == 1 ? v.y : (i >> 2) == 2 ? v.z : v.w;
    return (x >> ((i & 3) * 8)) & 0xFF;
}

/* Exactly n (0..16) bytes of v to ring position r, wrapping at the ring's
 * end: 8/4/2/1-byte stores, so no byte past n is touched. */
template <class W>
__device__ __forceinline__ void ring_put(W& w, uint32_t r, const uint4& v, uint32_t n) {
    if (n == 0) {
        return;
    }
    if (r + n <= W::kSize) {
        uint8_t* p = &w.ring[r];
        if (n == 16) {
            __builtin_memcpy(p, &v, 16);
            return;
        }
        uint64_t lo = ((uint64_t)v.y << 32) | v.x;
        const uint64_t hi = ((uint64_t)v.w << 32) | v.z;
        if (n & 8) {
            __builtin_memcpy(p, &lo, 8);
            p += 8;
            lo = hi;
        }
        if (n & 4) {
            const uint32_t q = (uint32_t)lo;
            __builtin_memcpy(p, &q, 4);
            p += 4;
            lo >>= 32;
        }
        if (n & 2) {
            const uint16_t q = (uint16_t)lo;
            __builtin_memcpy(p, &q, 2);
            p += 2;
            lo >>= 16;
        }
        if (n & 1) {
            *p = (uint8_t)lo;
        }
        return;
    }
    for (uint32_t i = 0; i < n; i++) {
        w.ring[(r + i) & (W::kSize - 1)] = (uint8_t)block_byte(v, i);
    }
}

/* Four selector bytes (r + i) mod p, i = 0..3, packed, for 3 <= p <= 15 and
 * r < p, given pb = p and cp = 0x80 - p in every byte. A bytewise compare
 * and subtract in one word: every byte stays below 128, so no carry or
 * borrow crosses a byte, and ge * 255 is (ge << 8) - ge modulo 2^32. */
__device__ __forceinline__ uint32_t period_sel(uint32_t r, uint32_t pb, uint32_t cp) {
    const uint32_t sv = __builtin_amdgcn_perm(0u, r, 0u) + 0x03020100u;
    const uint32_t ge = ((sv + cp) >> 7) & 0x01010101u;
    return sv - (((ge << 8) - ge) & pb);
}

/* Bytes of the 16-byte block S picked by four selector bytes (each 0..15). */
__device__ __forceinline__ uint32_t gather16(const uint4& S, uint32_t sel) {
    const uint32_t lo = __builtin_amdgcn_perm(S.y, S.x, sel & 0x07070707u);
    const uint32_t hi = __builtin_amdgcn_perm(S.w, S.z, sel & 0x07070707u);
    const uint32_t pick = 0x03020100u + ((sel >> 1) & 0x04040404u);
    return __builtin_amdgcn_perm(hi, lo, pick);
}

/* The source period of a lane match lies wholly in flushed history, so its
 * bytes can be loaded before the window's literals are placed. */
__device__ __forceinline__ bool period_in_history(uint64_t m, uint32_t off, uint32_t len,
                                                  uint64_t hist_lo, uint64_t cap) {
    const uint64_t y = m - off;
    const uint32_t need = len < off ? len : off;
    return y + need <= hist_lo && y + ((need + 15) & ~15u) <= cap;
}

/* Four 16-byte global loads of a history period of need (1..64) bytes at
 * `at`, issued together: block k starts 16 k bytes in, or at the period's
 * last block where the period is shorter, so every load stays inside the
 * period's 16-byte blocks. Straight-line on purpose: a load under a
 * condition of its own leaves the compiler copying its result where the
 * condition ends, and the copy waits for the load right there. */
struct Period {
    uint4 h0, h1, h2, h3;
};
__device__ __forceinline__ Period load_period(gcu8* at, uint32_t need) {
    const uint32_t last = (need - 1) & ~15u;
    Period p;
    p.h0 = gload16(at);
    p.h1 = gload16(at + (last < 16u ? last : 16u));
    p.h2 = gload16(at + (last < 32u ? last : 32u));
    p.h3 = gload16(at + last);
    return p;
}
/* What lanes with no history period load instead (never written). */
__device__ __align__(16) uint8_t no_period[16];

/* One lane copies a match of len (<= kLaneMatch) bytes at output position
 * m with offset off into the ring; every source byte precedes m or is
 * written earlier by this lane. Two steps:
 *  1. the period S = [m - off, m - off + min(len, off)) goes to the ring at
 *     m: from flushed history (from_history: hp holds it, loaded by the
 *     caller with up to four 16-byte loads issued together), or from the
 *     ring with independent 16-byte reads, or bytewise where it straddles
 *     the ring's floor or wraps;
 *  2. an overlapping match (off < len) repeats it: for off < 16 the first
 *     16 bytes are built in registers from S (byte k = S[k mod off]) and
 *     the rest is copied from D = off * floor(16 / off) bytes back, 8 or 16
 *     bytes a step; for off >= 16, 16 bytes a step from off back. Where the
 *     match wraps the ring, bytewise. */
template <class W>
__device__ __forceinline__ void lane_match_copy(W& w, gcu8* dst, uintptr_t da,
                                                uint64_t m, uint32_t off, uint32_t len,
                                                uint64_t hist_lo, uint64_t cap, bool from_history,
                                                const Period& hp) {
    const uint64_t y = m - off;
    const uint32_t need = len < off ? len : off;
    const uint32_t rd = ring_index<W::kSize>(da, m);
    const uint32_t ry = ring_index<W::kSize>(da, y);
    /* step 1 */
    if (from_history || (y >= hist_lo && ry + need <= W::kSize)) {
        uint4 h0, h1, h2, h3;
        if (from_history) {
            h0 = hp.h0;
            h1 = hp.h1;
            h2 = hp.h2;
            h3 = hp.h3;
        } else {
            /* reads past `need` (at most 15 bytes, still inside this
             * WaveLds) are never stored */
            __builtin_memcpy(&h0, &w.ring[ry], 16);
            h1 = h2 = h3 = make_uint4(0, 0, 0, 0);
            if (need > 16) {
                __builtin_memcpy(&h1, &w.ring[ry + 16], 16);
            }
            if (need > 32) {
                __builtin_memcpy(&h2, &w.ring[ry + 32], 16);
            }
            if (need > 48) {
                __builtin_memcpy(&h3, &w.ring[ry + 48], 16);
            }
        }
        if (rd + need <= W::kSize) {
            /* whole blocks, then one exact tail */
            const uint32_t full = need >> 4;
            if (full > 0) {
                __builtin_memcpy(&w.ring[rd], &h0, 16);
            }
            if (full > 1) {
                __builtin_memcpy(&w.ring[rd + 16], &h1, 16);
            }
            if (full > 2) {
                __builtin_memcpy(&w.ring[rd + 32], &h2, 16);
            }
            if (full > 3) {
                __builtin_memcpy(&w.ring[rd + 48], &h3, 16);
            }
            const uint4 t = full == 0 ? h0 : full == 1 ? h1 : full == 2 ? h2 : h3;
            ring_put(w, rd + 16 * full, t, need & 15);
        } else {
            ring_put(w, rd, h0, need < 16 ? need : 16);
            if (need > 16) {
                ring_put(w, (rd + 16) & (W::kSize - 1), h1, need - 16 < 16 ? need - 16 : 16);
            }
            if (need > 32) {
                ring_put(w, (rd + 32) & (W::kSize - 1), h2, need - 32 < 16 ? need - 32 : 16);
            }
            if (need > 48) {
                ring_put(w, (rd + 48) & (W::kSize - 1), h3, need - 48);
            }
        }
    } else {
        for (uint32_t i = 0; i < need; i++) {
            w.ring[(rd + i) & (W::kSize - 1)] = history_byte(w, dst, y + i, hist_lo);
        }
    }
    if (need == len) {
        return;
    }
    /* step 2: off < len */
    if (off < 16 && rd + (len > 16 ? len : 16) <= W::kSize) {
        uint4 S;
        __builtin_memcpy(&S, &w.ring[rd], 16);
        uint32_t p = off;
        if (p <= 2) {
            /* periods 1 and 2 as period 4 */
            S.x = __builtin_amdgcn_perm(0u, S.x, p == 1 ? 0x00000000u : 0x01000100u);
            p = 4;
        }
        const uint32_t pb = __builtin_amdgcn_perm(0u, p, 0u);
        const uint32_t cp = __builtin_amdgcn_perm(0u, 0x80u - p, 0u);
        uint32_t q[4];
        uint32_t r = 0;
#pragma unroll
        for (int t = 0; t < 4; t++) {
            q[t] = gather16(S, period_sel(r, pb, cp));
            r += 4;
            r = r >= p ? r - p : r;
            r = r >= p ? r - p : r;
        }
        ring_put(w, rd, make_uint4(q[0], q[1], q[2], q[3]), len < 16 ? len : 16);
        /* D - 1 for off = 1..15, a nibble each */
        const uint32_t D = (uint32_t)((0x0EDCBA98FDBEFEFFull >> (4 * (off - 1))) & 15) + 1;
        if (D == 16) {
            for (uint32_t i = 16; i < len; i += 16) {
                uint4 v;
                __builtin_memcpy(&v, &w.ring[rd + i - 16], 16);
                ring_put(w, rd + i, v, len - i < 16 ? len - i : 16);
            }
        } else {
            for (uint32_t i = 16; i < len; i += 8) {
                uint2 v;
                __builtin_memcpy(&v, &w.ring[rd + i - D], 8);
                ring_put(w, rd + i, make_uint4(v.x, v.y, 0, 0), len - i < 8 ? len - i : 8);
            }
        }
    } else if (off >= 16 && rd + len <= W::kSize) {
        for (uint32_t i = off; i < len; i += 16) {
            uint4 v;
            __builtin_memcpy(&v, &w.ring[rd + i - off], 16);
            ring_put(w, rd + i, v, len - i < 16 ? len - i : 16);
        }
    } else {
        for (uint32_t i = off; i < len; i++) {
            w.ring[(rd + i) & (W::kSize - 1)] = w.ring[(rd + i - off) & (W::kSize - 1)];
        }
    }
}

/* --------------------------------------------------- global slow path */

/* dst[0, n) = src[0, n), whole wave, global to global. Destination-aligned
 * 16-byte blocks; each lane's source read is an unaligned 16-byte load that
 * stays inside [src, src + n). */
__device__ __forceinline__ void wave_copy_global(gu8* dst, gcu8* src,
                                                 uint64_t n, uint32_t lane) {
    if (n == 0) {
        return;
    }
    const uintptr_t d0 = (uintptr_t)dst;
    uint64_t head = ((d0 + 15) & ~(uintptr_t)15) - d0;
    if (head > n) {
        head = n;
    }
    if (lane < head) {
        dst[lane] = src[lane];
    }
    const uint64_t body = (n - head) / 16;
    /* four loads a lane in flight before their stores: one wave keeps
     * 4 KiB moving, where a load-then-store step waits a full memory
     * latency per KiB (2000 random chunks of up to 4 MiB, each one literal
     * run: 5.0 ms that way, 1.05 ms like this; eight deep: 1.03 ms) */
    constexpr uint32_t kDeep = 4;
    uint64_t b = lane;
    for (; b + (kDeep - 1) * kWave < body; b += kDeep * kWave) {
        uint4 v[kDeep];
#pragma unroll
        for (uint32_t u = 0; u < kDeep; u++) {
            v[u] = gload16(src + head + 16 * (b + u * kWave));
        }
#pragma unroll
        for (uint32_t u = 0; u < kDeep; u++) {
            gstore16(dst + head + 16 * (b + u * kWave), v[u]);
        }
    }
    for (; b < body; b += kWave) {
        gstore16(dst + head + 16 * b, gload16(src + head + 16 * b));
    }
    const uint64_t done = head + 16 * body;
    if (lane < n - done) {
        dst[done + lane] = src[done + lane];
    }
    wave_order();
}

/* Largest period staged in the ring by wave_match_global. */
constexpr uint32_t kStagedPeriod = 1024;
static_assert(kStagedPeriod + 16 <= kRing, "a staged period fits the ring");

/* dst[m + i] = dst[m - off + (i mod off)] for i < len (cap bounds dst),
 * whole wave, in global memory; every source byte precedes m and is final.
 * The ring is free scratch here (everything in it is flushed):
 *  - off >= len: one contiguous copy;
 *  - off < kStagedPeriod: the period, extended by 16 bytes, is staged in
 *    the ring once, then every lane stores 16-byte slices read from it at
 *    (i mod off) -- one global round trip for the whole match, where a
 *    byte-per-lane gather pays one per 64 bytes and whole-period rounds one
 *    per period;
 *  - otherwise: rounds of off (>= kStagedPeriod) bytes. */
template <class W>
__device__ __forceinline__ void wave_match_global(W& w, gu8* dst, uint64_t m,
                                                  uint64_t off, uint64_t len, uint64_t cap,
                                                  uint32_t lane) {
    if (off >= len) {
        wave_copy_global(dst + m, dst + m - off, len, lane);
        return;
    }
    if (off >= kStagedPeriod) {
        for (uint64_t r0 = 0; r0 < len; r0 += off) {
            const uint64_t n = len - r0 < off ? len - r0 : off;
            wave_copy_global(dst + m + r0, dst + m + r0 - off, n, lane);
        }
        return;
    }
    const uint32_t o = (uint32_t)off;
    const uint64_t from = m - o;
    /* the period into ring[0, o): 16-byte loads where the block stays in
     * [0, cap); the bytes a block reads past the period are overwritten by
     * the extension below */
    for (uint32_t b = 16 * lane; b < o; b += 16 * kWave) {
        uint4 v;
        if (from + b + 16 <= cap) {
            v = gload16(dst + from + b);
        } else {
            uint32_t q[4] = {0, 0, 0, 0};
            for (uint32_t i = 0; i < 16 && b + i < o; i++) {
                q[i >> 2] |= (uint32_t)dst[from + b + i] << ((i & 3) * 8);
            }
            v = make_uint4(q[0], q[1], q[2], q[3]);
        }
        __builtin_memcpy(&w.ring[b], &v, 16);
    }
    wave_order();
    /* ring[o + t] = ring[t mod o] for t < 16 */
    if (lane < 16) {
        uint32_t k = lane;
        while (k >= o) {
            k -= o;
        }
        const uint8_t e = w.ring[k];
        w.ring[o + lane] = e;
    }
    wave_order();
    uint32_t r = (16 * lane) % o;
    const uint32_t step = (16 * kWave) % o;
    for (uint64_t b = 16 * lane; b < len; b += 16 * kWave) {
        uint4 v;
        __builtin_memcpy(&v, &w.ring[r], 16);
        if (len - b >= 16) {
            gstore16(dst + m + b, v);
        } else {
            const uint32_t n = (uint32_t)(len - b);
            for (uint32_t i = 0; i < n; i++) {
                dst[m + b + i] = (uint8_t)block_byte(v, i);
            }
        }
        r += step;
        r = r >= o ? r - o : r;
    }
    wave_order();
}

/* ------------------------------------------------- group execution */

/* Clears bit j of a wave-uniform mask in one scalar instruction. */
__device__ __forceinline__ void clear_bit(uint64_t& mask, uint32_t j) {
    asm("s_bitset0_b64 %0, %1" : "+s"(mask) : "s"(j));
}

/* 16 ring bytes from index r, wrapping at the ring's end. */
template <class W>
__device__ __forceinline__ uint4 ring_read16(const W& w, uint32_t r) {
    uint4 v;
    if (r <= W::kSize - 16) {
        __builtin_memcpy(&v, &w.ring[r], 16);
    } else {
        uint32_t q[4] = {0, 0, 0, 0};
        for (uint32_t i = 0; i < 16; i++) {
            q[i >> 2] |= (uint32_t)w.ring[(r + i) & (W::kSize - 1)] << ((i & 3) * 8);
        }
        v = make_uint4(q[0], q[1], q[2], q[3]);
    }
    return v;
}

/* Bytes t < keep of a, the rest of b. */
__device__ __forceinline__ uint4 take_bytes(const uint4& a, const uint4& b, uint32_t keep) {
    if (keep >= 16) {
        return a;
    }
    uint32_t m[4];
#pragma unroll
    for (uint32_t q = 0; q < 4; q++) {
        const uint32_t c = keep > 4 * q ? keep - 4 * q : 0;
        m[q] = c >= 4 ? 0xFFFFFFFFu : (1u << (8 * c)) - 1u;
    }
    return make_uint4((a.x & m[0]) | (b.x & ~m[0]), (a.y & m[1]) | (b.y & ~m[1]),
                      (a.z & m[2]) | (b.z & ~m[2]), (a.w & m[3]) | (b.w & ~m[3]));
}

/* x mod d for x < 4096 and d >= 1, by the float reciprocal: the quotient
 * x * (1 / d) never exceeds the true one (the reciprocal is within one
 * ulp) and its error stays below one part in 2^11, so it falls short by at
 * most one, and only at exact multiples; one correction suffices. */
__device__ __forceinline__ uint32_t small_mod(uint32_t x, uint32_t d) {
    const float inv = __builtin_amdgcn_rcpf((float)d);
    const uint32_t r = x - __umul24(d, (uint32_t)((float)x * inv));
    return r >= d ? r - d : r;
}

/* The match [m, m + len) with offset o, whole wave, 16 bytes a lane a round
 * of 1024, into the ring, when its period [m - o, m - o + min(o, len))
 * lies in the ring. Match byte i is period byte i mod o, so every lane
 * reads only the period, never bytes this match writes, and the lanes are
 * independent. Lane k's block starts at match byte i0 = 16 k (+ 1024 a
 * round), at period phase s = i0 mod o:
 *  - o >= 16: the block is period bytes s, s + 1, ..., wrapping at most
 *    once, i.e. ring bytes from m - o + s while t < o - s, and from one
 *    period further back, m - 2 o + s, after that: two 16-byte reads and a
 *    byte mask;
 *  - o < 16: the first 16 ring bytes from m - o hold the period, and byte
 *    t is its byte (s + t) mod o, gathered with v_perm selectors (periods
 *    1 and 2 are widened to 4 first).
 * Each block stores exactly its bytes of the match. */
template <class W>
__device__ __forceinline__ void ring_match_wide(W& w, uintptr_t da, uint64_t m, uint32_t o,
                                                uint32_t len, uint32_t lane) {
    const uint32_t rm = ring_index<W::kSize>(da, m);
    const uint32_t rp = ring_index<W::kSize>(da, m - o);
    uint32_t s = small_mod(16 * lane, o);
    const uint32_t step = small_mod(16 * kWave, o);
    if (o >= 16) {
        for (uint32_t i0 = 16 * lane; i0 < len; i0 += 16 * kWave) {
            const uint4 a = ring_read16(w, (rp + s) & (W::kSize - 1));
            const uint4 b = ring_read16(w, (rp + s - o) & (W::kSize - 1));
            const uint32_t n = len - i0 < 16 ? len - i0 : 16;
            ring_put(w, (rm + i0) & (W::kSize - 1), take_bytes(a, b, o - s), n);
            s += step;
            s = s >= o ? s - o : s;
        }
        return;
    }
    uint4 S = ring_read16(w, rp);
    uint32_t p = o;
    if (p <= 2) {
        /* periods 1 and 2 as period 4 */
        S.x = __builtin_amdgcn_perm(0u, S.x, p == 1 ? 0x00000000u : 0x01000100u);
        p = 4;
    }
    const uint32_t pb = __builtin_amdgcn_perm(0u, p, 0u);
    const uint32_t cp = __builtin_amdgcn_perm(0u, 0x80u - p, 0u);
    for (uint32_t i0 = 16 * lane; i0 < len; i0 += 16 * kWave) {
        uint32_t q[4];
        uint32_t r = s;
#pragma unroll
        for (int t = 0; t < 4; t++) {
            q[t] = gather16(S, period_sel(r, pb, cp));
            r += 4;
            r = r >= p ? r - p : r;
            r = r >= p ? r - p : r;
        }
        const uint32_t n = len - i0 < 16 ? len - i0 : 16;
        ring_put(w, (rm + i0) & (W::kSize - 1), make_uint4(q[0], q[1], q[2], q[3]), n);
        s += step;
        s = s >= o ? s - o : s;
    }
}

/* The members of one window that a group executes, as the copy phase needs
 * them: per lane, the member's literal start (window-relative), literal
 * and match lengths, offset and output offset from the group start; and,
 * uniform, how many members run in full, whether member `take` adds its
 * terminal literal only, and the group's output length. */
struct Group {
    uint32_t lit, lit_len, match_len, offset, drel;
    uint32_t take;
    bool terminal;
    uint32_t end_rel;
};

/* One lane's literal of lit_len (<= kLaneLit) bytes, staged at cw + lit,
 * to the ring at output position lit_at: 16 bytes a step, the last step
 * exact (a Snappy literal has no match after it to absorb an overrun). */
template <class W>
__device__ __forceinline__ void put_literal(W& w, const uint8_t* cw, uint32_t lit,
                                            uint32_t lit_len, uintptr_t da, uint64_t lit_at) {
    uint32_t i0 = 0;
    for (; i0 + 16 <= lit_len; i0 += 16) {
        uint4 v;
        __builtin_memcpy(&v, cw + lit + i0, 16);
        const uint32_t r = ring_index<W::kSize>(da, lit_at + i0);
        if (r <= W::kSize - 16) {
            __builtin_memcpy(&w.ring[r], &v, 16);
        } else {
            for (uint32_t b = 0; b < 16; b++) {
                w.ring[(r + b) & (W::kSize - 1)] = (uint8_t)block_byte(v, b);
            }
        }
    }
    if (i0 < lit_len) {
        uint4 v;
        __builtin_memcpy(&v, cw + lit + i0, 16);
        ring_put(w, ring_index<W::kSize>(da, lit_at + i0), v, lit_len - i0);
    }
}

/* Executes group g of the window staged at cw into the ring at output
 * positions [dp, dp + g.end_rel), flushes whole 16-byte blocks (all of them
 * when the chunk is done) and advances dp. */
template <class W>
__device__ __forceinline__ void execute_group(W& w, const uint8_t* cw, const Group& g,
                                              gu8* dst, uint64_t cap, uint64_t& dp,
                                              uint64_t& flushed, uint64_t ring_lo, bool done,
                                              uint32_t lane) {
    const uintptr_t da = (uintptr_t)dst;
    const uint32_t take = g.take;
    const bool full = lane < take;
    const bool active = full || (g.terminal && lane == take);
    const uint32_t lit = g.lit, lit_len = g.lit_len, match_len = g.match_len,
                   offset = g.offset, drel = g.drel;
    const uint64_t D0 = dp;
    const uint64_t D1 = dp + g.end_rel;
    const uint64_t ring_floor = D1 + 16 > W::kSize ? D1 + 16 - W::kSize : 0;
    const uint64_t hist_lo = ring_lo > ring_floor ? ring_lo : ring_floor;
    const uint64_t lit_at = D0 + drel;
    const uint64_t m_at = lit_at + lit_len;
    /* a literal the whole wave copies in stream order */
    const bool wide_lit = lit_len > kLaneLit;
    /* matches that read only history or their own literals go by their own
     * lane; the history ones load their period now, so the load overlaps
     * the literal phase */
    const uint64_t y = m_at - offset;
    const uint64_t need = match_len < offset ? match_len : offset;
    const bool independent = y + need <= D0 || y >= lit_at;
    const bool lane_match = full && !wide_lit && match_len <= kLaneMatch && independent;
    const bool from_history =
        lane_match && period_in_history(m_at, offset, match_len, hist_lo, cap);
    /* every lane loads, with no branch around the loads (whose results
     * would then be copied, waiting for them here); a lane without a
     * history period reads the zero block no_period */
    const Period hp = load_period(from_history ? dst + y : (gcu8*)no_period,
                                  from_history ? (uint32_t)need : 1u);
    /* phase 1: the lanes' literals */
    if (active && !wide_lit) {
        put_literal(w, cw, lit, lit_len, da, lit_at);
    }
    wave_order();
    /* phase 2: the lane matches */
    if (lane_match) {
        lane_match_copy(w, dst, da, m_at, offset, match_len, hist_lo, cap, from_history, hp);
    }
    wave_order();
    /* phase 3: the rest, in stream order, whole wave. The common member, a
     * short match (<= 64 bytes) whose period is in the ring, is described by
     * its lane up front in two words, so the loop pays two v_readlane and a
     * byte a lane for it: fast_a = ring index of the match (13 bits), of
     * its period (13 bits), length - 1 (6 bits); fast_b = o = min(offset,
     * 64) in bits 0-7 and ceil(4096 / o) above them. Byte i copies period
     * byte i mod offset, which precedes the match, so the lanes do not
     * depend on each other. For i <= 64 the quotient floor(i / o) is
     * (i * ceil(4096 / o)) >> 12: with ceil(4096 / o) = (4096 + e) / o,
     * 0 <= e < o, the scaled product is i / o + i * e / (4096 o), and
     * i * e < 64 * 64 keeps the excess below the 1 / o that separates i / o
     * from the next integer. Two 24-bit multiplies and a shift replace a
     * reciprocal and its correction. */
    const bool rest_fast = full && !lane_match && !wide_lit && match_len != 0 &&
                           match_len <= kWave && y >= hist_lo;
    static_assert(W::kSize <= 8192, "ring indices are packed in 13 bits");
    const uint32_t fast_a = ring_index<W::kSize>(da, m_at) | ring_index<W::kSize>(da, y) << 13 |
                            (match_len - 1) << 26;
    const uint32_t fast_o = offset - 1 < kWave ? offset : kWave;
    const uint32_t fast_b = fast_o | ((4096u + fast_o - 1) / fast_o) << 8;
    /* in stream order: runs of short ring matches (one straight-line step
     * each, in a loop with no other branch), each run ended by the next
     * other member */
    uint64_t fast = __ballot(rest_fast);
    uint64_t other = __ballot(active && (wide_lit || (full && !lane_match))) & ~fast;
    for (;;) {
        /* the fast members before the next other one */
        uint64_t run = other != 0 ? fast & ((other & (0 - other)) - 1) : fast;
        fast &= ~run;
        while (run != 0) {
            const uint32_t j = __builtin_ctzll(run);
            clear_bit(run, j);
            const uint32_t ja = __builtin_amdgcn_readlane(fast_a, j);
            const uint32_t jb = __builtin_amdgcn_readlane(fast_b, j);
            const uint32_t jo = jb & 255;
            const uint32_t rm = ja & 8191;
            const uint32_t rs = (ja >> 13) & 8191;
            const uint32_t jm = (ja >> 26) + 1;
            /* lane mod jo (lanes below jo get a zero quotient); lanes past
             * the match write their own sink byte, so no lane is masked */
            const uint32_t r = lane - __umul24(__umul24(lane, jb >> 8) >> 12, jo);
            uint8_t* at = lane < jm ? &w.ring[(rm + lane) & (W::kSize - 1)] : &w.sink[lane];
            *at = w.ring[(rs + r) & (W::kSize - 1)];
            wave_order();
        }
        if (other == 0) {
            break;
        }
        /* a wide literal, then the match unless the member
         * is the terminal literal (no rest member is a lane match) */
        const uint32_t j = __builtin_ctzll(other);
        clear_bit(other, j);
        const uint32_t jl = __builtin_amdgcn_readlane(lit_len, j);
        const uint32_t jlit = __builtin_amdgcn_readlane(lit, j);
        const uint32_t jd = __builtin_amdgcn_readlane(drel, j);
        const uint64_t jd0 = D0 + jd;
        if (jl > kLaneLit) {
            for (uint32_t i = lane; i < jl; i += kWave) {
                w.ring[ring_index<W::kSize>(da, jd0 + i)] = cw[jlit + i];
            }
        }
        wave_order();
        if (j < take) {
            const uint32_t jm = __builtin_amdgcn_readlane(match_len, j);
            const uint32_t joff = __builtin_amdgcn_readlane(offset, j);
            const uint64_t jm0 = jd0 + jl;
            const uint64_t jfrom = jm0 - joff;
            /* byte i copies period byte i mod joff, kept per lane as i
             * steps by 64; lane mod joff and 64 mod joff as in the fast
             * loop (for joff > 64 they are lane and 64) */
            const uint32_t jb = __builtin_amdgcn_readlane(fast_b, j);
            const uint32_t jo = jb & 255;
            const uint32_t jq = jb >> 8;
            uint32_t r = lane - __umul24(__umul24(lane, jq) >> 12, jo);
            const uint32_t step =
                joff > kWave ? kWave : kWave - __umul24(__umul24(kWave, jq) >> 12, jo);
            for (uint32_t i = lane; i < jm; i += kWave) {
                w.ring[ring_index<W::kSize>(da, jm0 + i)] =
                    history_byte(w, dst, jfrom + r, hist_lo);
                r += step;
                r = r >= joff ? r - joff : r;
            }
        }
        wave_order();
    }
    dp = D1;
    const uintptr_t a_end = done ? da + dp : ((da + dp) & ~(uintptr_t)15);
    const uint64_t to = a_end > da + flushed ? (uint64_t)(a_end - da) : flushed;
    if (to > flushed) {
        flush_ring(w, dst, flushed, to, lane);
        flushed = to;
    }
}

/* ------------------------------------------------- window parse */

/* Follows the links from `start` (a scalar chain: each step reads the next
 * position out of the lane that holds the current one). Lane j of `member`
 * receives member j's position. The walk ends at a stop link (that position
 * is not a member), past the window, or after 64 members; it returns the
 * position after the last member and the member count. */
template <uint32_t kC, uint32_t kMinStep, bool kStraight>
__device__ __forceinline__ uint32_t walk_chain_regions(const uint32_t (&link)[kC],
                                                       uint32_t start, uint32_t& count_out,
                                                       uint32_t& member) {
    static_assert(kStop >= kC * kWave + kWave, "a stop link must leave every region");
    static_assert(kMinStep >= 2, "links go forward");
    /* most members one region can hold, and the count at or below which
     * the walk takes the unchecked loop there */
    constexpr uint32_t kRegionMax = (kWave - 1) / kMinStep + 1;
    constexpr uint32_t kFastMax = kWave - kRegionMax;
    /* A straight run of guarded regions with no early exit (an exit edge
     * out of the unrolled loop has the compiler re-materialise live values
     * at every region), whatever kStraight says. A region is entered only
     * while pos lies in it and fewer than 64 members are held; a stop link
     * lies past every region, so the walk's end needs no flag of its own
     * and is settled once, after the last region. */
    uint32_t pos = start, count = 0, prev = kStop;
#pragma unroll
    for (uint32_t k = 0; k < kC; k++) {
        uint32_t rel = pos - k * kWave;
        if ((rel | count) >= kWave) {
            continue; /* pos lies in a later region, or the walk has ended */
        }
        /* Per member: read the link of pos out of its lane, record pos as
         * member `count`, step. The loop leaves once pos is outside this
         * region (a stop link lands far outside) or 64 members are held.
         * Only scalar registers are written apart from `member`. Two
         * members a turn, the current position alternating between pos and
         * nx, so that no step moves it.
         * Links only go forward, by kMinStep bytes or more (an LZ4
         * sequence takes 3, a Snappy element 2), so a region holds at most
         * kRegionMax members and "pos < base + 64" is the whole region
         * check. Entered with at most kFastMax members, the walk cannot
         * pass 64 inside the region and takes the fast loop: the position itself selects
         * the lane (a wave64 v_readlane uses bits [5:0] of the select), and
         * a member costs a compare, a branch and the count. Otherwise the
         * checked loop tests the count as well. In the fast loop four
         * instructions separate a v_readlane from the one that uses its
         * result as the select, the 4 wait states that use needs. */
        uint32_t nx, t;
        /* The count is an operand bound to m0 ("{m0}"), so the compiler
         * moves it in and out and keeps its own uses of m0 apart:
         * v_writelane takes its value and lane select from scalar
         * registers, and only m0 as the lane select keeps it within one
         * constant-bus read. The leading s_nop covers a lane select (rel)
         * that the compiler may have just written from a vector
         * instruction, and m0 written just before its first use. */
        asm volatile(
            "s_nop 4\n\t"
            "s_cmp_gt_u32 %[cnt], %[fmax]\n\t"
            "s_cbranch_scc1 walk_loop_%=\n"
            "walk_fast_%=:\n\t"
            "v_readlane_b32 %[nx], %[lk], %[pos]\n\t"
            "v_writelane_b32 %[mem], %[pos], m0\n\t"
            "s_add_u32 m0, m0, 1\n\t"
            "s_cmp_lt_u32 %[nx], %[lim]\n\t"
            "s_cbranch_scc0 walk_out_nx_%=\n\t"
            "v_readlane_b32 %[pos], %[lk], %[nx]\n\t"
            "v_writelane_b32 %[mem], %[nx], m0\n\t"
            "s_add_u32 m0, m0, 1\n\t"
            "s_cmp_lt_u32 %[pos], %[lim]\n\t"
            "s_cbranch_scc1 walk_fast_%=\n\t"
            "s_mov_b32 %[prev], %[nx]\n\t"
            "s_branch walk_done_%=\n"
            "walk_loop_%=:\n\t"
            "v_readlane_b32 %[nx], %[lk], %[rel]\n\t"
            "v_writelane_b32 %[mem], %[pos], m0\n\t"
            "s_sub_u32 %[rel], %[nx], %[base]\n\t"
            "s_add_u32 m0, m0, 1\n\t"
            "s_or_b32 %[t], %[rel], m0\n\t"
            "s_cmp_lt_u32 %[t], 64\n\t"
            "s_cbranch_scc0 walk_out_nx_%=\n\t"
            "v_readlane_b32 %[pos], %[lk], %[rel]\n\t"
            "v_writelane_b32 %[mem], %[nx], m0\n\t"
            "s_sub_u32 %[rel], %[pos], %[base]\n\t"
            "s_add_u32 m0, m0, 1\n\t"
            "s_or_b32 %[t], %[rel], m0\n\t"
            "s_cmp_lt_u32 %[t], 64\n\t"
            "s_cbranch_scc1 walk_loop_%=\n\t"
            "s_mov_b32 %[prev], %[nx]\n\t"
            "s_branch walk_done_%=\n"
            "walk_out_nx_%=:\n\t"
            "s_mov_b32 %[prev], %[pos]\n\t"
            "s_mov_b32 %[pos], %[nx]\n"
            "walk_done_%=:"
            : [mem] "+v"(member), [pos] "+s"(pos), [cnt] "+{m0}"(count), [rel] "+s"(rel),
              [prev] "=&s"(prev), [nx] "=&s"(nx), [t] "=&s"(t)
            : [lk] "v"(link[k]), [base] "s"(k * kWave), [lim] "s"(k * kWave + kWave),
              [fmax] "s"(kFastMax)
            : "scc");
    }
    /* a stop link: the member that read it (prev, the last member) is no
     * member, and the walk ends there; a walk that started on a stop link
     * holds no member and stays as it is */
    if (pos == kStop && count != 0) {
        pos = prev;
        count -= 1;
    }
    count_out = uniform32(count);
    return uniform32(pos);
}

/* One region k of the walk (positions [64 k, 64 k + 64), links in lk<k>):
 * per member, read the link of pos out of its lane, record pos as member
 * m0, step; two members a turn, the current position alternating between
 * pos and nx so that no step moves it. Links only go forward by kMinStep
 * bytes or more, so a region holds at most kRegionMax members and
 * "pos < 64 k + 64" is the whole region check; entered with at most
 * kFastMax members the walk cannot pass 64 inside the region and takes the
 * fast loop (the position itself selects the lane: a wave64 v_readlane
 * uses bits [5:0] of the select), else the checked loop, which tests the
 * count as well. Four instructions separate a v_readlane from the one that
 * uses its result as the select (the 4 wait states that use needs), and
 * three separate an m0 update from the v_writelane that reads it. A
 * region is entered only while pos lies in it and fewer than 64 members are
 * held ((pos - base) | count below 64); leaving it (pos = the next position)
 * falls into the next region's check. A stop link lies past every region,
 * so the walk's end is settled once, in WALK_END. */
#define WALK_REGION(k, B, L, N)                                                             \
    "r" #k "_%=:\n\t"                                                                       \
    "s_sub_u32 %[rel], %[pos], " #B "\n\t"                                                  \
    "s_or_b32 %[t], %[rel], m0\n\t"                                                         \
    "s_cmp_lt_u32 %[t], 64\n\t"                                                             \
    "s_cbranch_scc0 r" #N "_%=\n\t"                                                         \
    "s_cmp_gt_u32 m0, %[fmax]\n\t"                                                          \
    "s_cbranch_scc1 c" #k "_%=\n"                                                           \
    "f" #k "_%=:\n\t"                                                                       \
    "v_readlane_b32 %[nx], %[lk" #k "], %[pos]\n\t"                                         \
    "v_writelane_b32 %[mem], %[pos], m0\n\t"                                                \
    "s_add_u32 m0, m0, 1\n\t"                                                               \
    "s_cmp_lt_u32 %[nx], " #L "\n\t"                                                        \
    "s_cbranch_scc0 o" #k "_%=\n\t"                                                         \
    "v_readlane_b32 %[pos], %[lk" #k "], %[nx]\n\t"                                         \
    "v_writelane_b32 %[mem], %[nx], m0\n\t"                                                 \
    "s_add_u32 m0, m0, 1\n\t"                                                               \
    "s_cmp_lt_u32 %[pos], " #L "\n\t"                                                       \
    "s_cbranch_scc1 f" #k "_%=\n\t"                                                         \
    "s_mov_b32 %[prev], %[nx]\n\t"                                                          \
    "s_branch x" #k "_%=\n"                                                                 \
    "c" #k "_%=:\n\t"                                                                       \
    "v_readlane_b32 %[nx], %[lk" #k "], %[rel]\n\t"                                         \
    "v_writelane_b32 %[mem], %[pos], m0\n\t"                                                \
    "s_sub_u32 %[rel], %[nx], " #B "\n\t"                                                   \
    "s_add_u32 m0, m0, 1\n\t"                                                               \
    "s_or_b32 %[t], %[rel], m0\n\t"                                                         \
    "s_cmp_lt_u32 %[t], 64\n\t"                                                             \
    "s_cbranch_scc0 o" #k "_%=\n\t"                                                         \
    "v_readlane_b32 %[pos], %[lk" #k "], %[rel]\n\t"                                        \
    "v_writelane_b32 %[mem], %[nx], m0\n\t"                                                 \
    "s_sub_u32 %[rel], %[pos], " #B "\n\t"                                                  \
    "s_add_u32 m0, m0, 1\n\t"                                                               \
    "s_or_b32 %[t], %[rel], m0\n\t"                                                         \
    "s_cmp_lt_u32 %[t], 64\n\t"                                                             \
    "s_cbranch_scc1 c" #k "_%=\n\t"                                                         \
    "s_mov_b32 %[prev], %[nx]\n\t"                                                          \
    "s_branch x" #k "_%=\n"                                                                 \
    "o" #k "_%=:\n\t"                                                                       \
    "s_mov_b32 %[prev], %[pos]\n\t"                                                         \
    "s_mov_b32 %[pos], %[nx]\n"                                                             \
    "x" #k "_%=:\n"

/* The walk's entry and exits around the regions: the count lives in m0
 * during the walk (saved and restored around it): v_writelane takes its
 * value and lane select from scalar registers, and only m0 as the lane
 * select keeps it within one constant-bus read. The leading s_nop covers a
 * lane select the compiler may have just written from a vector instruction,
 * and m0 written just before its first use. Past the last region: a stop
 * link (pos = 0xffff) read by the last member (prev) makes it no member, and
 * the walk ends there (a walk that started on a stop link holds none);
 * otherwise the walk ended past the window or at 64 members. */
#define WALK_BEGIN                                                                          \
    "s_mov_b32 %[m0s], m0\n\t"                                                              \
    "s_mov_b32 m0, %[cnt]\n\t"                                                              \
    "s_nop 4\n"
#define WALK_END(N)                                                                         \
    "r" #N "_%=:\n\t"                                                                       \
    "s_cmpk_eq_u32 %[pos], 0xffff\n\t"                                                      \
    "s_cbranch_scc0 end_%=\n\t"                                                             \
    "s_cmp_eq_u32 m0, 0\n\t"                                                                \
    "s_cbranch_scc1 end_%=\n\t"                                                             \
    "s_mov_b32 %[pos], %[prev]\n\t"                                                         \
    "s_sub_u32 m0, m0, 1\n"                                                                 \
    "end_%=:\n\t"                                                                           \
    "s_mov_b32 %[cnt], m0\n\t"                                                              \
    "s_mov_b32 m0, %[m0s]"
#define WALK_OPERANDS                                                                       \
    : [mem] "+v"(member), [pos] "+s"(pos), [cnt] "+s"(count), [rel] "=&s"(rel),             \
      [prev] "+s"(prev), [nx] "=&s"(nx), [t] "=&s"(t), [m0s] "=&s"(m0_saved)

/* The walk of walk_chain_regions, with the same results: for windows of up
 * to eight regions as one asm block (WALK_REGION per region), else region
 * by region. */
template <uint32_t kC, uint32_t kMinStep, bool kStraight>
__device__ __forceinline__ uint32_t walk_chain(const uint32_t (&link)[kC], uint32_t start,
                                               uint32_t& count_out, uint32_t& member) {
    static_assert(kStop >= kC * kWave + kWave, "a stop link must leave every region");
    static_assert(kMinStep >= 2, "links go forward");
    (void)kStraight;
    /* most members one region can hold, and the count at or below which
     * the walk takes the unchecked loop there */
    constexpr uint32_t kRegionMax = (kWave - 1) / kMinStep + 1;
    constexpr uint32_t kFastMax = kWave - kRegionMax;
    if constexpr (kC > 8) {
        /* sixteen regions in one block keep all sixteen link registers live
         * at once and spill (5% slower on LZ4 than a block per region) */
        return walk_chain_regions<kC, kMinStep, kStraight>(link, start, count_out, member);
    } else {
        uint32_t pos = start, count = 0, rel, prev = kStop, nx, t, m0_saved;
        /* one asm block for the whole walk: no compiler code between regions
         * (Snappy's seven regions: 2% faster than a block per region) */
        if constexpr (kC == 8) {
            asm volatile(WALK_BEGIN WALK_REGION(0, 0, 64, 1) WALK_REGION(1, 64, 128, 2)
                             WALK_REGION(2, 128, 192, 3) WALK_REGION(3, 192, 256, 4)
                                 WALK_REGION(4, 256, 320, 5) WALK_REGION(5, 320, 384, 6)
                                     WALK_REGION(6, 384, 448, 7) WALK_REGION(7, 448, 512, 8)
                                         WALK_END(8)
                         WALK_OPERANDS
                         : [lk0] "v"(link[0]), [lk1] "v"(link[1]), [lk2] "v"(link[2]),
                           [lk3] "v"(link[3]), [lk4] "v"(link[4]), [lk5] "v"(link[5]),
                           [lk6] "v"(link[6]), [lk7] "v"(link[7]), [fmax] "s"(kFastMax)
                         : "scc");
        } else if constexpr (kC == 7) {
            asm volatile(WALK_BEGIN WALK_REGION(0, 0, 64, 1) WALK_REGION(1, 64, 128, 2)
                             WALK_REGION(2, 128, 192, 3) WALK_REGION(3, 192, 256, 4)
                                 WALK_REGION(4, 256, 320, 5) WALK_REGION(5, 320, 384, 6)
                                     WALK_REGION(6, 384, 448, 7) WALK_END(7)
                         WALK_OPERANDS
                         : [lk0] "v"(link[0]), [lk1] "v"(link[1]), [lk2] "v"(link[2]),
                           [lk3] "v"(link[3]), [lk4] "v"(link[4]), [lk5] "v"(link[5]),
                           [lk6] "v"(link[6]), [fmax] "s"(kFastMax)
                         : "scc");
        } else {
            static_assert(kC == 7, "walk regions are spelled out for these widths");
        }
        count_out = uniform32(count);
        return uniform32(pos);
    }
}
#undef WALK_REGION
#undef WALK_BEGIN
#undef WALK_END
#undef WALK_OPERANDS

/* The chain from `start`, as walk_chain decides it (members while in the
 * window and under 64; a position with a stop link is no member), by
 * pointer doubling over the window's links in LDS (jump, kC * 64 entries)
 * instead of one dependent step per member. Round r (r = 0..5) has the
 * table map a position to the one 2^r links on: lanes [2^r, 2^(r+1)) take
 * their member from lane (lane - 2^r) with one lookup, then every entry
 * looks itself up to double the reach. A link at or past the window end
 * (kStop included) maps to itself. The rounds stop once the last member
 * found is past the window. `link` is clobbered. */
template <uint32_t kC>
__device__ __forceinline__ uint32_t walk_doubling(uint32_t (&link)[kC], uint32_t start,
                                                  uint16_t* jump, uint32_t& count_out,
                                                  uint32_t& member, uint32_t lane) {
    constexpr uint32_t kW = kC * kWave;
    static_assert(kStop >= kW && kStop <= 0xFFFFu, "a stop link is past every window");
#pragma unroll
    for (uint32_t k = 0; k < kC; k++) {
        jump[k * kWave + lane] = (uint16_t)link[k];
    }
    wave_order();
    uint32_t m = start; /* lane i: member i's position once found */
    uint32_t h = 1;
    for (;;) {
        const uint32_t from = (uint32_t)__shfl((int)m, (int)((lane - h) & (kWave - 1)));
        if (lane >= h && lane < 2 * h) {
            m = from < kW ? jump[from] : from;
        }
        const uint32_t last = uniform32(__builtin_amdgcn_readlane(m, 2 * h - 1));
        if (last >= kW || h == kWave / 2) {
            break;
        }
#pragma unroll
        for (uint32_t k = 0; k < kC; k++) {
            link[k] = link[k] < kW ? jump[link[k]] : link[k];
        }
        wave_order();
#pragma unroll
        for (uint32_t k = 0; k < kC; k++) {
            jump[k * kWave + lane] = (uint16_t)link[k];
        }
        wave_order();
        h *= 2;
    }
    const uint64_t found = 2 * h >= kWave ? ~0ull : (1ull << (2 * h)) - 1;
    const uint64_t past = __ballot(m >= kW) & found;
    uint32_t count, pos;
    if (past != 0) {
        const uint32_t s = (uint32_t)__builtin_ctzll(past);
        pos = __builtin_amdgcn_readlane(m, s);
        count = s;
        if (pos == kStop && s != 0) {
            count = s - 1;
            pos = __builtin_amdgcn_readlane(m, s - 1);
        }
    } else {
        /* 64 members in the window: the table maps 32 links on */
        const uint32_t m64 = jump[uniform32(__builtin_amdgcn_readlane(m, kWave / 2))];
        count = kWave;
        pos = m64;
        if (m64 == kStop) {
            count = kWave - 1;
            pos = __builtin_amdgcn_readlane(m, kWave - 1);
        }
    }
    member = m;
    count_out = uniform32(count);
    return uniform32(pos);
}

/* What parsing one window decided. kGroup: execute `g` (staged at `delta`
 * in the stage buffer) and advance the source by `consumed`. kSlow: one
 * sequence the window cannot hold, by the scalar ladder (`q`, with the
 * cursor after it in sp/dp). kFail: the chunk ends with `status`. */
enum ParseKind : uint32_t { kParseGroup = 0, kParseSlow = 1, kParseFail = 2 };

struct Parsed {
    uint32_t kind;
    int status;
    bool done;        /* the chunk ends after this group / sequence */
    uint32_t delta;
    Group g;
    Lz4Step q;        /* kSlow: the sequence; kGroup with lead: runs first */
    bool lead;        /* kGroup: q precedes the group */
    uint64_t lead_dp; /* output position after q */
};

/* Stages the window at sp into `stage`, speculates, walks and resolves it
 * (src/lz4_block.h:95-258 for every member, in parallel) and advances
 * sp and dp past what it decided. Reads the source only. kDoubling: walk
 * the chain by pointer doubling over `jump` (the one-wave form, where the
 * wave's own latency is the pace), else serially (the two-wave parser,
 * which leaves LDS and issue to the copier; `jump` is unused). */
template <uint32_t kC, bool kDoubling>
__device__ __forceinline__ Parsed parse_window(uint8_t* stage, uint16_t* jump, gcu8* src,
                                               uint64_t src_size, uint64_t cap, uint64_t& sp,
                                               uint64_t& dp, uint32_t lane,
                                               uint32_t& touch) {
    Parsed out;
    out.kind = kParseFail;
    out.status = kOk;
    out.done = false;
    out.lead = false;
    out.lead_dp = 0;
    const uint64_t base = sp; /* window start; member positions are relative to it */
    out.delta = stage_into<kC>(stage, src, src_size, sp, lane);
    touch = touch_ahead<kC>(src, src_size, sp, lane);
    const uint8_t* cw = stage + out.delta;
    /* the scalar ladder reads a sequence's header from the stage while it
     * lies there (an LDS round trip per byte instead of a global one) */
    const StagedBytes staged{src, cw, base, stage_bytes(kC) + 16 - out.delta};

    /* speculate: the link of each candidate position */
    uint32_t link[kC];
#pragma unroll
    for (uint32_t k = 0; k < kC; k++) {
        const Cand c = parse_candidate(cw, k * kWave + lane);
        link[k] = c.complex ? kStop : c.next;
    }

    /* a window that opens on a sequence it cannot hold: the exact scalar
     * ladder decides that one, and the chain goes on from just past it
     * (the speculated links stay valid) */
    uint32_t start = 0;
    if (uniform32(__builtin_amdgcn_readlane(link[0], 0)) == kStop) {
        Lz4Cursor cur{src, src_size, cap, sp, dp};
        bool last = false;
        const int st = (int)uniform32((uint32_t)lz4_next<true>(cur, out.q, last, staged));
        if (st != kOk) {
            out.status = st;
            return out;
        }
        sp = uniform64(cur.sp);
        dp = uniform64(cur.dp);
        out.lead = true;
        out.lead_dp = dp;
        const uint64_t adv = sp - base;
        if (uniform32(last) != 0 || adv >= kC * kWave) {
            out.kind = kParseSlow;
            out.lead = false;
            out.done = uniform32(last) != 0;
            return out;
        }
        start = (uint32_t)adv;
    }

    /* walk the true chain */
    uint32_t count = 0;
    uint32_t member = 0; /* lane j: position of member j */
    const uint32_t pos = kDoubling
                             ? walk_doubling<kC>(link, start, jump, count, member, lane)
                             : walk_chain<kC, 3, false>(link, start, count, member);

    bool slow = count == 0;
    if (!slow) {
        /* resolve members: the exact ladder for each member's sequence */
        const bool is_member = lane < count;
        const Cand c = parse_candidate(cw, is_member ? member : 0);
        const uint32_t out_len = is_member ? c.lit_len + c.match_len : 0;
        const uint32_t incl = wave_inclusive_sum(out_len, lane);
        const uint32_t drel = incl - out_len;
        const uint64_t d0 = dp + drel; /* absolute destination */
        int ev = kOk;
        bool terminal = false;
        if (is_member) {
            uint64_t s = base + member;  /* absolute source position */
            if (s >= src_size) {
                ev = kCorrupt;
            } else {
                s++;
                if ((c.tok >> 4) == 15) {
                    if (s + 15 >= src_size) {
                        ev = kCorrupt;
                    } else if (15 > cap - d0) {
                        ev = kTooSmall;
                    } else {
                        s++;
                        if (s + 15 > src_size) {
                            ev = kCorrupt;
                        }
                    }
                }
                const uint64_t L = c.lit_len;
                if (ev == kOk) {
                    if (L > src_size - s) {
                        ev = kCorrupt;
                    } else if (L > cap - d0) {
                        ev = kTooSmall;
                    } else if (cap - d0 - L < kLz4SlackDst || src_size - s - L < kLz4SlackSrc) {
                        if (s + L != src_size) {
                            ev = kCorrupt;
                        } else if (cap == 0 && c.tok != 0) {
                            ev = kCorrupt;
                        } else {
                            terminal = true;
                        }
                    } else {
                        const uint64_t dl = d0 + L;
                        uint64_t M = (c.tok & 15) + 4;
                        s += L + 2;
                        if (c.offset == 0 || c.offset > dl) {
                            ev = kCorrupt;
                        } else {
                            if ((c.tok & 15) == 15) {
                                if (s >= src_size) {
                                    ev = kCorrupt;
                                } else if (M > cap - dl) {
                                    ev = kTooSmall;
                                } else {
                                    s++;
                                    M = c.match_len;
                                    if (s + kLz4MatchMargin > src_size) {
                                        ev = kCorrupt;
                                    }
                                }
                            }
                            if (ev == kOk) {
                                if (M > cap - dl) {
                                    ev = kTooSmall;
                                } else if (M > cap - dl - kLz4LastLiterals) {
                                    ev = kCorrupt;
                                }
                            }
                        }
                    }
                }
            }
        }
        const uint32_t mine = terminal ? c.lit_len : out_len;
        const bool cut = is_member && drel + mine > batch_max(kDoubling);
        const bool flag = is_member && (cut || ev != kOk || terminal);
        const uint64_t flags = __ballot(flag);
        uint32_t take = count; /* members executed in full */
        bool take_terminal = false;
        int status = kOk;
        if (flags != 0) {
            const uint32_t j = __builtin_ctzll(flags);
            take = j;
            const bool j_cut = __builtin_amdgcn_readlane((uint32_t)cut, j) != 0;
            const int j_ev = (int)__builtin_amdgcn_readlane((uint32_t)ev, j);
            if (j_cut) {
                slow = j == 0;
            } else if (j_ev != kOk) {
                status = j_ev;
            } else {
                take_terminal = true;
            }
        }
        take = uniform32(take);
        take_terminal = uniform32(take_terminal) != 0;
        slow = uniform32(slow) != 0;
        status = (int)uniform32((uint32_t)status);
        if (status != kOk) {
            out.status = status;
            return out;
        }
        if (!slow && (take != 0 || take_terminal)) {
            const uint32_t last = take_terminal ? take : take - 1;
            out.kind = kParseGroup;
            out.g.lit = c.lit;
            out.g.lit_len = c.lit_len;
            out.g.match_len = c.match_len;
            out.g.offset = c.offset;
            out.g.drel = drel;
            out.g.take = take;
            out.g.terminal = take_terminal;
            out.g.end_rel = uniform32(__builtin_amdgcn_readlane(incl, last) -
                                      (take_terminal ? __builtin_amdgcn_readlane(c.match_len, last)
                                                     : 0));
            const uint64_t next_sp =
                take_terminal ? src_size
                              : base + (take == count ? pos : __builtin_amdgcn_readlane(member, take));
            sp = uniform64(next_sp);
            dp = uniform64(dp + out.g.end_rel);
            out.done = take_terminal;
            return out;
        }
    }
    if (out.lead) {
        /* nothing after the leading sequence fits this window */
        out.kind = kParseSlow;
        out.lead = false;
        return out;
    }
    /* one sequence by the exact scalar ladder */
    Lz4Cursor cur{src, src_size, cap, sp, dp};
    bool last = false;
    const int st = (int)uniform32((uint32_t)lz4_next<true>(cur, out.q, last, staged));
    if (st != kOk) {
        out.status = st;
        return out;
    }
    out.kind = kParseSlow;
    out.done = uniform32(last) != 0;
    sp = uniform64(cur.sp);
    dp = uniform64(cur.dp);
    return out;
}

/* ------------------------------------------------- Snappy window parse */

/* Longest literal a Snappy window resolves; longer ones take the scalar
 * ladder. A window literal therefore ends below kC * 64 + 5 + 256. */
constexpr uint32_t kSnappyWinLit = 256;
static_assert(kCandSplit * 64 + 5 + kSnappyWinLit + 16 <= stage_bytes(kCandSplit) + 16 &&
                  kCandPipe * 64 + 5 + kSnappyWinLit + 16 <= stage_bytes(kCandPipe) + 16,
              "stage covers a Snappy window");

/* The element a tag at window position c would encode
 * (src/snappy_block.h:288-435): a literal (lit, lit_len; match_len 0) or a
 * copy-1/2/4 (match_len, offset; lit_len 0). One 8-byte read at the tag. */
__device__ __forceinline__ Cand snappy_element(const uint8_t* cw, uint32_t c) {
    Cand k;
    uint64_t h;
    __builtin_memcpy(&h, cw + c, 8);
    const uint32_t tag = (uint32_t)h & 0xFF;
    k.tok = tag;
    k.complex = false;
    if ((tag & 3) == 0) {
        const uint32_t len6 = tag >> 2;
        uint32_t header = 1;
        uint64_t length = (uint64_t)len6 + 1;
        if (len6 > 59) {
            const uint32_t extra = len6 - 59;
            header += extra;
            const uint64_t bytes = h >> 8;
            length = (extra == 4 ? (bytes & 0xFFFFFFFFull) : (bytes & ((1ull << (8 * extra)) - 1))) + 1;
        }
        if (length > kSnappyWinLit) {
            k.complex = true;
            length = 0;
        }
        k.lit = c + header;
        k.lit_len = (uint32_t)length;
        k.match_len = 0;
        k.offset = 0;
        k.next = c + header + (uint32_t)length;
    } else {
        const uint32_t kind = tag & 3;
        const uint32_t ob = kind == 1 ? 1 : (kind == 2 ? 2 : 4);
        uint32_t offset = (uint32_t)(h >> 8);
        offset &= ob == 1 ? 0xFFu : (ob == 2 ? 0xFFFFu : 0xFFFFFFFFu);
        uint32_t length;
        if (kind == 1) {
            offset |= (tag >> 5) << 8;
            length = 4 + ((tag >> 2) & 7u);
        } else {
            length = (tag >> 2) + 1;
        }
        k.lit = c + 1 + ob;
        k.lit_len = 0;
        k.match_len = length;
        k.offset = offset;
        k.next = c + 1 + ob;
    }
    return k;
}

/* The member a tag at window position c would start: one element, or a
 * literal element fused with the copy element right after it (the pair an
 * LZ4 sequence carries as literals + match), so a window's 64 members hold
 * more of Snappy's short elements. A copy is fused only while its tag lies
 * before the source end (lim = source bytes from the window start). */
__device__ __forceinline__ Cand parse_candidate_snappy(const uint8_t* cw, uint32_t c,
                                                      uint32_t lim) {
    Cand k = snappy_element(cw, c);
    if (k.match_len == 0 && !k.complex && k.next < lim) {
        const Cand m = snappy_element(cw, k.next);
        if (m.match_len != 0) {
            k.match_len = m.match_len;
            k.offset = m.offset;
            k.next = m.next;
        }
    }
    return k;
}

/* As parse_window, for Snappy elements after the preamble: every member's
 * rungs run in parallel (every reject is CORRUPT_INPUT, so the first
 * failing member in stream order decides), and the member whose position
 * is exactly the end of the source ends the chunk when the output so far
 * equals the declared length. kDoubling as for parse_window. */
template <uint32_t kC, bool kDoubling>
__device__ __forceinline__ Parsed parse_window_snappy(uint8_t* stage, uint16_t* jump, gcu8* src,
                                                      uint64_t src_size, uint64_t cap,
                                                      uint64_t declared, uint64_t& sp,
                                                      uint64_t& dp, uint32_t lane,
                                                      uint32_t& touch) {
    Parsed out;
    out.kind = kParseFail;
    out.status = kOk;
    out.done = false;
    out.lead = false;
    out.lead_dp = 0;
    const uint64_t base = sp;
    out.delta = stage_into<kC>(stage, src, src_size, sp, lane);
    touch = touch_ahead<kC>(src, src_size, sp, lane);
    const uint8_t* cw = stage + out.delta;
    /* source bytes from the window start, clamped: a fused copy must start
     * before the source end */
    const uint32_t lim = src_size - base >= 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)(src_size - base);

    uint32_t link[kC];
#pragma unroll
    for (uint32_t k = 0; k < kC; k++) {
        const Cand c = parse_candidate_snappy(cw, k * kWave + lane, lim);
        link[k] = c.complex ? kStop : c.next;
    }
    /* a window that opens on a literal it cannot hold: the scalar ladder
     * decides it and the chain goes on from just past it */
    uint32_t start = 0;
    if (uniform32(__builtin_amdgcn_readlane(link[0], 0)) == kStop) {
        SnappyCursor cur{src, src_size, cap, sp, dp, declared};
        SnappyStep e;
        bool last = false;
        const int st = (int)uniform32((uint32_t)snappy_next(cur, e, last));
        if (st != kOk) {
            out.status = st;
            return out;
        }
        out.q.lit_src = e.is_copy ? 0 : e.from;
        out.q.lit_dst = e.to;
        out.q.lit_len = e.is_copy ? 0 : e.length;
        out.q.match_dst = e.to + (e.is_copy ? 0 : e.length);
        out.q.match_len = e.is_copy ? e.length : 0;
        out.q.offset = e.is_copy ? (uint32_t)e.offset : 1u;
        sp = uniform64(cur.sp);
        dp = uniform64(cur.dp);
        out.lead = true;
        out.lead_dp = dp;
        const uint64_t adv = sp - base;
        if (uniform32(last) != 0 || adv >= kC * kWave) {
            out.kind = kParseSlow;
            out.lead = false;
            out.done = uniform32(last) != 0;
            return out;
        }
        start = (uint32_t)adv;
    }
    uint32_t count = 0;
    uint32_t member = 0;
    const uint32_t pos = kDoubling
                             ? walk_doubling<kC>(link, start, jump, count, member, lane)
                             : walk_chain<kC, 2, true>(link, start, count, member);

    bool slow = count == 0;
    if (!slow) {
        const bool is_member = lane < count;
        const Cand c = parse_candidate_snappy(cw, is_member ? member : 0, lim);
        const uint64_t s = base + member;
        const bool at_end = is_member && s >= src_size;
        const uint32_t out_len = is_member && !at_end ? c.lit_len + c.match_len : 0;
        const uint32_t incl = wave_inclusive_sum(out_len, lane);
        const uint32_t drel = incl - out_len;
        const uint64_t d0 = dp + drel;
        int ev = kOk;
        if (is_member) {
            if (at_end) {
                ev = s == src_size && d0 == declared ? kOk : kCorrupt;
            } else if (d0 > declared) {
                ev = kCorrupt;
            } else if (c.lit_len != 0 || c.match_len == 0) {
                /* a literal element (src/snappy_block.h:345-370), then, in a
                 * fused member, the copy element after it at (s2, d2) */
                const uint64_t header = c.lit - member;
                if (header > src_size - s) {
                    ev = kCorrupt;
                } else if (c.lit_len > src_size - s - header) {
                    ev = kCorrupt;
                } else if (c.lit_len > declared - d0) {
                    ev = kCorrupt;
                } else if (c.match_len != 0) {
                    const uint64_t s2 = s + header + c.lit_len;
                    const uint64_t d2 = d0 + c.lit_len;
                    const uint64_t header2 = c.next - (c.lit + c.lit_len);
                    if (header2 > src_size - s2) {
                        ev = kCorrupt;
                    } else if (c.offset == 0 || (uint64_t)c.offset > d2) {
                        ev = kCorrupt;
                    } else if (c.match_len > declared - d2) {
                        ev = kCorrupt;
                    }
                }
            } else {
                /* a copy element (src/snappy_block.h:371-435) */
                const uint64_t header = c.lit - member;
                if (header > src_size - s) {
                    ev = kCorrupt;
                } else if (c.offset == 0 || (uint64_t)c.offset > d0) {
                    ev = kCorrupt;
                } else if (c.match_len > declared - d0) {
                    ev = kCorrupt;
                }
            }
        }
        const bool cut = is_member && !at_end && drel + out_len > batch_max(kDoubling);
        const bool flag = is_member && (cut || ev != kOk || at_end);
        const uint64_t flags = __ballot(flag);
        uint32_t take = count;
        bool ended = false;
        int status = kOk;
        if (flags != 0) {
            const uint32_t j = __builtin_ctzll(flags);
            take = j;
            const bool j_cut = __builtin_amdgcn_readlane((uint32_t)cut, j) != 0;
            const int j_ev = (int)__builtin_amdgcn_readlane((uint32_t)ev, j);
            if (j_ev != kOk) {
                status = j_ev;
            } else if (j_cut) {
                slow = j == 0;
            } else {
                ended = true;
            }
        }
        take = uniform32(take);
        ended = uniform32(ended) != 0;
        slow = uniform32(slow) != 0;
        status = (int)uniform32((uint32_t)status);
        if (status != kOk) {
            out.status = status;
            return out;
        }
        if (!slow && (take != 0 || ended)) {
            out.kind = kParseGroup;
            out.g.lit = c.lit;
            out.g.lit_len = c.lit_len;
            out.g.match_len = c.match_len;
            out.g.offset = c.offset;
            out.g.drel = drel;
            out.g.take = take;
            out.g.terminal = false;
            out.g.end_rel = take == 0 ? 0u : uniform32(__builtin_amdgcn_readlane(incl, take - 1));
            const uint32_t consumed = take == count ? pos : __builtin_amdgcn_readlane(member, take);
            sp = uniform64(ended ? src_size : base + consumed);
            dp = uniform64(dp + out.g.end_rel);
            out.done = ended;
            return out;
        }
    }
    if (out.lead) {
        out.kind = kParseSlow;
        out.lead = false;
        return out;
    }
    /* one element by the exact scalar ladder */
    SnappyCursor cur{src, src_size, cap, sp, dp, declared};
    SnappyStep e;
    bool last = false;
    const int st = (int)uniform32((uint32_t)snappy_next(cur, e, last));
    if (st != kOk) {
        out.status = st;
        return out;
    }
    out.kind = kParseSlow;
    out.done = uniform32(last) != 0;
    out.q.lit_src = e.is_copy ? 0 : e.from;
    out.q.lit_dst = e.to;
    out.q.lit_len = e.is_copy ? 0 : e.length;
    out.q.match_dst = e.to + (e.is_copy ? 0 : e.length);
    out.q.match_len = e.is_copy ? e.length : 0;
    out.q.offset = e.is_copy ? (uint32_t)e.offset : 1u;
    sp = uniform64(cur.sp);
    dp = uniform64(cur.dp);
    return out;
}

/* A window of either format. */
template <int kFormat, uint32_t kC, bool kDoubling>
__device__ __forceinline__ Parsed parse_any(uint8_t* stage, uint16_t* jump, gcu8* src,
                                            uint64_t src_size,
                                            uint64_t cap, uint64_t declared, uint64_t& sp,
                                            uint64_t& dp, uint32_t lane,
                                            uint32_t& touch) {
    if (kFormat == kFormatSnappy) {
        return parse_window_snappy<kC, kDoubling>(stage, jump, src, src_size, cap, declared, sp,
                                                  dp, lane, touch);
    }
    return parse_window<kC, kDoubling>(stage, jump, src, src_size, cap, sp, dp, lane, touch);
}

/* Start of a chunk: the Snappy preamble and its declared length against the
 * capacity (src/snappy_block.h:190-263); nothing for LZ4. */
template <int kFormat>
__device__ __forceinline__ int chunk_begin(gcu8* src, uint64_t src_size, uint64_t cap,
                                           uint64_t& sp, uint64_t& declared) {
    sp = 0;
    declared = 0;
    if (kFormat != kFormatSnappy) {
        return kOk;
    }
    SnappyCursor c{src, src_size, cap, 0, 0, 0};
    const int st = (int)uniform32((uint32_t)snappy_begin(c));
    sp = uniform64(c.sp);
    declared = uniform64(c.declared);
    return st;
}

/* Largest offset whose period a slow match takes from the ring. */
constexpr uint32_t kSlowRingPeriod = 1024;

/* Executes a sequence the window could not hold (a length with two or more
 * extension bytes, a literal past the window). A literal of at most 64
 * bytes goes into the ring a byte a lane; a longer one runs global to
 * global after the ring is flushed up to it. A match whose period (offset
 * <= kSlowRingPeriod) is in the ring is expanded there 16 bytes a lane
 * (ring_match_wide), in pieces that never lap unflushed bytes or the
 * period, with the finished blocks flushed before each piece; any other
 * match runs global to global (wave_match_global), after everything before
 * it is flushed. On return fewer than 16 bytes are unflushed (none when
 * `done`: the chunk ends here). */
template <class W>
__device__ __forceinline__ void execute_slow(W& w, gcu8* src, gu8* dst,
                                             uint64_t cap, const Lz4Step& q, uint64_t dp_after,
                                             uint64_t& dp,
                                             uint64_t& flushed, uint64_t& ring_lo, bool done,
                                             uint32_t lane) {
    const uintptr_t da = (uintptr_t)dst;
    if (q.lit_len <= kWave) {
        if (lane < q.lit_len) {
            w.ring[ring_index<W::kSize>(da, q.lit_dst + lane)] = src[q.lit_src + lane];
        }
        wave_order();
    } else {
        flush_ring(w, dst, flushed, q.lit_dst, lane);
        wave_copy_global(dst + q.lit_dst, src + q.lit_src, q.lit_len, lane);
        flushed = q.lit_dst + q.lit_len;
        ring_lo = flushed; /* the ring holds none of it */
    }
    if (q.match_len != 0) {
        uint64_t m = q.match_dst;
        const uint32_t o = q.offset;
        const uint64_t floor = m + 16 > W::kSize ? m + 16 - W::kSize : 0;
        if (o <= kSlowRingPeriod && m - o >= (ring_lo > floor ? ring_lo : floor)) {
            /* a piece leaves the period and the (< 16) unflushed bytes
             * below m in place */
            constexpr uint32_t kPiece = W::kSize - kSlowRingPeriod - 32;
            uint64_t left = q.match_len;
            while (left != 0) {
                /* whole blocks below m (compared as addresses: near the
                 * chunk's start the aligned address can precede dst) */
                const uintptr_t a_to = (da + m) & ~(uintptr_t)15;
                if (a_to > da + flushed) {
                    flush_ring(w, dst, flushed, (uint64_t)(a_to - da), lane);
                    flushed = (uint64_t)(a_to - da);
                }
                const uint32_t piece = left < kPiece ? (uint32_t)left : kPiece;
                ring_match_wide(w, da, m, o, piece, lane);
                wave_order();
                m += piece;
                left -= piece;
            }
        } else {
            flush_ring(w, dst, flushed, m, lane);
            wave_match_global(w, dst, m, o, q.match_len, cap, lane);
            flushed = m + q.match_len;
            ring_lo = flushed; /* wave_match_global used the ring as scratch */
        }
    }
    /* leave fewer than 16 unflushed bytes (none when the chunk is done):
     * the next group's output may fill the ring up to its batch limit */
    dp = dp_after;
    const uintptr_t a_end = done ? da + dp : ((da + dp) & ~(uintptr_t)15);
    if (a_end > da + flushed) {
        flush_ring(w, dst, flushed, (uint64_t)(a_end - da), lane);
        flushed = (uint64_t)(a_end - da);
    }
}

/* ------------------------------------------------- two-wave pipeline */

/* The window record the parsing wave hands the copying wave: one Parsed,
 * laid out per lane in LDS. */
struct alignas(16) PipeRec {
    uint16_t lit[kWave];
    uint16_t lit_len[kWave];
    uint16_t match_len[kWave];
    uint32_t offset[kWave];
    uint16_t drel[kWave];
    uint32_t kind;
    int32_t status;
    uint32_t done;
    uint32_t delta;
    uint32_t take;
    uint32_t terminal;
    uint32_t end_rel;
    uint32_t lead;              /* kGroup: the slow sequence below runs first */
    uint64_t lit_src, lit_dst, lit_len_s, match_dst, offset_s, match_len_s, dp_after;
};

/* A block's LDS. The two-wave form keeps kS window slots (a stage buffer
 * and a record each; slot 0's stage sits in the ring's LDS); the parser
 * fills them in turn and the copier runs them in the same order. */
template <uint32_t kS>
struct alignas(16) PipeLds {
    union {
        struct {
            PipeRing w;                        /* copier's ring, slot 0's stage */
            uint8_t stage[kS - 1][kStageBuf];  /* the other slots' stages */
            PipeRec rec[kS];
        } pipe;                                /* two-wave form */
        struct {
            WaveLds a, b;                      /* one wave each, once the block splits */
        } split;
    };
    uint32_t chunk_idx;
    uint32_t fin;           /* lockstep: the step whose record ends the chunk */
    uint32_t pub[kS];       /* slotted: 1 + sequence number of each slot's record */
    uint32_t freed[kS];     /* slotted: 1 + sequence number of the record run last there */
    uint32_t broken;        /* slotted: a wait ran out, and no wave waits any more */

    __device__ uint8_t* stage_of(uint32_t slot) {
        return slot == 0 ? pipe.w.stage : pipe.stage[slot - 1];
    }
};

/* Slotted blocks signal through LDS words, each written by one wave and
 * polled by the other (s_sleep between reads). A wait longer than any
 * window could take (2^27 polls: seconds) marks the block broken, and a
 * broken block waits no more (false). */
typedef __attribute__((address_space(3))) volatile uint32_t lds_flag;
__device__ __forceinline__ bool await_flag(const lds_flag* flag, uint32_t want, lds_flag* broken) {
    for (uint32_t i = 0;; i++) {
        if (uniform32(*flag) == want) {
            break;
        }
        if (uniform32(*broken) != 0) {
            return false;
        }
        if (i == (1u << 27)) {
            *broken = 1;
            return false;
        }
        __builtin_amdgcn_s_sleep(1);
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    return true;
}
__device__ __forceinline__ void raise_flag(lds_flag* flag, uint32_t value) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    *flag = value;
}

static_assert(kLaneLit < 65536 && kBatchMaxPipe < 65536 && kStageBuf < 65536,
              "window records hold 16-bit fields");

/* ------------------------------------------------------------- LZ4 */

struct ChunkOut {
    int status;
    uint64_t produced;
};

/* One wave per chunk: parse a window, then execute it, in turn.
 * kParseOnly: parse only (statuses and sizes, no output bytes). */
template <int kFormat, bool kParseOnly>
__device__ __forceinline__ ChunkOut decode_chunk(WaveLds& w, gcu8* src,
                                                 uint64_t src_size, gu8* dst, uint64_t cap,
                                                 uint32_t lane) {
    uint64_t sp = 0, declared = 0, pdp = 0, dp = 0, flushed = 0, ring_lo = 0;
    uint64_t fuel = src_size == UINT64_MAX ? UINT64_MAX : src_size + 2;
    int status = chunk_begin<kFormat>(src, src_size, cap, sp, declared);
    while (status == kOk) {
        if (fuel-- == 0) {
            status = kCorrupt;
            break;
        }
        uint32_t touch = 0;
        const Parsed pr =
            parse_any<kFormat, kCandSplit, true>(w.stage, w.jump, src, src_size, cap, declared,
                                                 sp, pdp, lane, touch);
        if (pr.kind == kParseFail) {
            status = pr.status;
            break;
        }
        if (kParseOnly) {
            dp = pdp;
        } else if (pr.kind == kParseGroup) {
            if (pr.lead) {
                execute_slow(w, src, dst, cap, pr.q, pr.lead_dp, dp, flushed, ring_lo, false,
                             lane);
            }
            execute_group(w, w.stage + pr.delta, pr.g, dst, cap, dp, flushed, ring_lo, pr.done,
                          lane);
        } else {
            execute_slow(w, src, dst, cap, pr.q, pdp, dp, flushed, ring_lo, pr.done, lane);
        }
        /* the look-ahead loads land by now */
        asm volatile("" ::"v"(touch));
        if (pr.done) {
            break;
        }
    }
    ChunkOut r;
    r.status = status;
    r.produced = dp;
    return r;
}

/* The parser's side of one window of the two-wave form: parse (or hand
 * over the chunk's failure) into record r and stage buffer `stage`;
 * returns whether this record ends the chunk. */
template <int kFormat>
__device__ __forceinline__ bool parse_into(PipeRec& r, uint8_t* stage, int begun, uint64_t& fuel,
                                           gcu8* src, uint64_t src_size, uint64_t cap,
                                           uint64_t declared, uint64_t& sp, uint64_t& dp,
                                           uint32_t lane, uint32_t& touch) {
    Parsed pr;
    if (begun != kOk || fuel-- == 0) {
        pr.kind = kParseFail;
        pr.status = begun != kOk ? begun : kCorrupt;
        pr.done = false;
        pr.delta = 0;
    } else {
        constexpr uint32_t kC = kFormat == kFormatSnappy ? kCandPipeSnappy : kCandPipe;
        pr = parse_any<kFormat, kC, false>(stage, nullptr, src, src_size, cap, declared, sp, dp,
                                           lane, touch);
    }
    r.kind = pr.kind;
    r.status = pr.status;
    r.done = pr.done ? 1u : 0u;
    r.lead = pr.lead ? 1u : 0u;
    if (pr.kind == kParseSlow || pr.lead) {
        r.lit_src = pr.q.lit_src;
        r.lit_dst = pr.q.lit_dst;
        r.lit_len_s = pr.q.lit_len;
        r.match_dst = pr.q.match_dst;
        r.offset_s = pr.q.offset;
        r.match_len_s = pr.q.match_len;
        r.dp_after = pr.lead ? pr.lead_dp : dp;
    }
    if (pr.kind == kParseGroup) {
        r.lit[lane] = (uint16_t)pr.g.lit;
        r.lit_len[lane] = (uint16_t)pr.g.lit_len;
        r.match_len[lane] = (uint16_t)pr.g.match_len;
        r.offset[lane] = pr.g.offset;
        r.drel[lane] = (uint16_t)pr.g.drel;
        r.delta = pr.delta;
        r.take = pr.g.take;
        r.terminal = pr.g.terminal ? 1u : 0u;
        r.end_rel = pr.g.end_rel;
    }
    return pr.kind == kParseFail || pr.done;
}

/* The copier's side: run record r, whose window is staged in `stage`
 * (kParseOnly: only the output size advances); a failed record sets
 * status. Returns whether this record ends the chunk. */
template <bool kParseOnly>
__device__ __forceinline__ bool run_record(PipeRing& w, const PipeRec& r, const uint8_t* stage,
                                           gcu8* src, gu8* dst, uint64_t cap, uint64_t& dp,
                                           uint64_t& flushed, uint64_t& ring_lo, int& status,
                                           uint32_t lane) {
    /* every field read at once, before the kind decides which are used
     * (one LDS round trip; a slow or failed record's group fields are
     * stale and unused) */
    Group g;
    g.lit = r.lit[lane];
    g.lit_len = r.lit_len[lane];
    g.match_len = r.match_len[lane];
    g.offset = r.offset[lane];
    g.drel = r.drel[lane];
    g.take = uniform32(r.take);
    g.terminal = uniform32(r.terminal) != 0;
    g.end_rel = uniform32(r.end_rel);
    const uint32_t delta = uniform32(r.delta);
    const uint32_t kind = uniform32(r.kind);
    const bool done = uniform32(r.done) != 0;
    const bool lead = uniform32(r.lead) != 0;
    if (kind == kParseFail) {
        status = (int)uniform32((uint32_t)r.status);
        return true;
    }
    if (kParseOnly) {
        /* the parse ceiling: the records' sizes, no bytes */
        if (kind == kParseSlow || lead) {
            dp = uniform64(r.dp_after);
        }
        if (kind == kParseGroup) {
            dp += g.end_rel;
        }
        return done;
    }
    if (kind == kParseSlow || lead) {
        Lz4Step s;
        s.lit_src = uniform64(r.lit_src);
        s.lit_dst = uniform64(r.lit_dst);
        s.lit_len = uniform64(r.lit_len_s);
        s.match_dst = uniform64(r.match_dst);
        s.offset = (uint32_t)uniform64(r.offset_s);
        s.match_len = uniform64(r.match_len_s);
        execute_slow(w, src, dst, cap, s, uniform64(r.dp_after), dp, flushed, ring_lo,
                     kind == kParseSlow && done, lane);
    }
    if (kind == kParseGroup) {
        execute_group(w, stage + delta, g, dst, cap, dp, flushed, ring_lo, done, lane);
    }
    return done;
}

/* ------------------------------------------------------------ kernels */

__global__ void queue_reset(Queue* q) {
    const uint32_t t = threadIdx.x;
    if (t < kKeys) {
        q->bin_count[t] = 0;
        q->bin_cursor[t] = 0;
    }
    if (t < kSubQueues) {
        q->next[t * kSubStride] = 0;
    }
}

/* Order-preserving key of a cost below 2^63: the octave and the next
 * kOctaveBits bits below its leading one (costs below 2^kOctaveBits are
 * their own key). */
__host__ __device__ inline uint32_t cost_key(uint64_t cost) {
    if (cost < (1u << kOctaveBits)) {
        return (uint32_t)cost;
    }
    const uint32_t lg = 63u - (uint32_t)__builtin_clzll(cost);
    return lg << kOctaveBits |
           (uint32_t)((cost >> (lg - kOctaveBits)) & ((1u << kOctaveBits) - 1));
}

/* Bin of a chunk by its decode cost, compressed size plus capacity: a
 * chunk's time follows the bytes it produces (a 4 MiB run of short-offset
 * matches compresses to a few hundred KiB yet decodes as long as 4 MiB of
 * literals and far matches). Four bins per octave; larger cost, smaller
 * bin, decoded first. */
__device__ __forceinline__ uint32_t cost_bin(uint64_t src_size, uint64_t cap) {
    const uint64_t lim = 1ull << 61;
    const uint64_t a = src_size < lim ? src_size : lim;
    const uint64_t b = cap < lim ? cap : lim;
    /* the output, plus what matches add to it (capacity - compressed):
     * match bytes cost the decode its time, literal bytes are a copy */
    const uint64_t cost = b + (b > a ? b - a : 0);
    return kBins - 1 - cost_key(cost);
}

/* Queue key of a chunk: its class under the plan, then its bin. */
__device__ __forceinline__ uint32_t queue_key(const Plan& plan, uint64_t src_size, uint64_t cap) {
    const uint32_t bin = cost_bin(src_size, cap);
    return plan_two_wave(plan, bin) ? bin : kBins + bin;
}

/* Chunks per cost bin. A block counts its chunks in LDS and adds one
 * global atomic per bin it holds: a batch of like chunks (one bin) would
 * otherwise serialise one atomic per chunk on that bin's counter. */
__global__ void queue_count(const size_t* sizes, const size_t* caps, size_t n, Queue* q,
                            Plan plan) {
    __shared__ uint32_t local[kKeys];
    for (uint32_t t = threadIdx.x; t < kKeys; t += blockDim.x) {
        local[t] = 0;
    }
    __syncthreads();
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        atomicAdd(&local[queue_key(plan, sizes[i], caps[i])], 1u);
    }
    __syncthreads();
    for (uint32_t t = threadIdx.x; t < kKeys; t += blockDim.x) {
        if (local[t] != 0) {
            atomicAdd(&q->bin_count[t], local[t]);
        }
    }
}

__global__ void queue_scan(Queue* q) {
    if (threadIdx.x == 0) {
        uint32_t acc = 0;
        for (int b = 0; b < kKeys; b++) {
            q->bin_base[b] = acc;
            acc += q->bin_count[b];
        }
    }
}

/* Each chunk's place in the order: its bin's base, the block's range in
 * that bin (one global atomic per bin the block holds) and its rank in the
 * block (an LDS atomic). */
__global__ void queue_fill(const size_t* sizes, const size_t* caps, size_t n, Queue* q,
                           uint32_t* order, Plan plan) {
    __shared__ uint32_t local[kKeys];
    __shared__ uint32_t range[kKeys];
    for (uint32_t t = threadIdx.x; t < kKeys; t += blockDim.x) {
        local[t] = 0;
    }
    __syncthreads();
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t b = 0, rank = 0;
    if (i < n) {
        b = queue_key(plan, sizes[i], caps[i]);
        rank = atomicAdd(&local[b], 1u);
    }
    __syncthreads();
    for (uint32_t t = threadIdx.x; t < kKeys; t += blockDim.x) {
        if (local[t] != 0) {
            range[t] = atomicAdd(&q->bin_cursor[t], local[t]);
        }
    }
    __syncthreads();
    if (i < n) {
        order[q->bin_base[b] + range[b] + rank] = (uint32_t)i;
    }
}

/* The next queue position for this wave from kSubs sub-queues (1: the
 * single counter), or at least n once every one is drained (n < 2^32).
 * Every lane takes part in the atomic and only lane 0 adds: a
 * lane-0-only branch at a persistent loop's top and bottom lets the
 * compiler fold both into an exec-masked inner loop that only lane 0 can
 * leave. */
template <uint32_t kSubs>
__device__ __forceinline__ uint32_t queue_pop(Queue* q, size_t n, QueueCursor& c,
                                              uint32_t lane) {
    if (kSubs == 1) {
        return uniform32(atomicAdd(&q->next[0], lane == 0 ? 1u : 0u));
    }
    while (c.left != 0) {
        const uint32_t i =
            uniform32(atomicAdd(&q->next[c.sub * kSubStride], lane == 0 ? 1u : 0u));
        const uint64_t pos = (uint64_t)c.sub + (uint64_t)kSubs * i;
        if (pos < n) {
            return (uint32_t)pos;
        }
        c.sub = (c.sub + 1) % kSubs;
        c.left--;
    }
    return (uint32_t)n;
}

constexpr int kPipeBlock = 2 * kWave;
constexpr uint32_t kNoFin = 0xFFFFFFFFu;

/* Two waves per chunk. Wave 0 parses windows into the block's window
 * slots (a record and a stage buffer each) while wave 1 executes the
 * records in the same order. Parsing reads only the source and decides
 * every status; execution only writes the ring and the destination, so the
 * chunk's critical path is the longer of the two per window rather than
 * their sum. Slotted blocks (four slots) hand windows over through LDS
 * flags, so the parser may run three windows ahead; lockstep blocks (two
 * slots) take one barrier per window. Control flow around the barriers is
 * uniform for the block: both waves read the chunk index (and, lockstep,
 * the final step) from LDS after a barrier, and a barrier opens every
 * chunk so that no control word is rewritten while the other wave may
 * still read it. */
/* kWaves: waves per SIMD the register budget is cut for. 4 (128 VGPRs) for
 * launches that pipeline chunks; 5 (96 VGPRs, some spills on the two-wave
 * path) for launches where every chunk runs on one wave, which are
 * throughput-bound and gain from a fifth wave. kSubs: decode sub-queues,
 * kSubQueues for launches of many chunks per wave, else 1 (the single
 * counter, whose code keeps the fewest scalar registers live). */
template <int kFormat, int kWaves, uint32_t kSubs, bool kParseOnly>
__global__ void __launch_bounds__(kPipeBlock) __attribute__((amdgpu_waves_per_eu(kWaves)))
    piped_batch(const void* const* srcs, const size_t* src_sizes, void* const* dsts,
                const size_t* caps, size_t n, cudec_chunk_result* results, Queue* q,
                const uint32_t* order, Plan plan) {
    /* Slotted (four waves per SIMD): four window slots signalled through
     * LDS flags, so the parser runs up to three windows ahead and neither
     * wave waits on the other's slowest windows (a window's parse and copy
     * times vary independently). Lockstep (five waves per SIMD, where ten
     * blocks must fit a CU's LDS): two slots and a barrier per window. */
    constexpr bool kSlotted = kWaves == 4;
    constexpr uint32_t kS = kSlotted ? 4 : 2;
    __shared__ PipeLds<kS> L;
    const uint32_t lane = __lane_id();
    const bool parser = uniform32(threadIdx.x / kWave) == 0;
    uint32_t split_idx = 0xFFFFFFFFu; /* first chunk decoded one wave per chunk */
    QueueCursor cursor{blockIdx.x % kSubs, kSubs};
    if (kSlotted && parser && lane < kS) {
        L.pub[lane] = 0;
        L.freed[lane] = 0;
        L.broken = 0;
    }
    uint32_t seq = 0; /* slotted: windows of this block so far, counted alike by both waves */
    for (;;) {
        /* both waves are done reading the previous chunk's control words
         * (chunk_idx, fin) before the parser writes the next chunk's */
        __syncthreads();
        if (parser) {
            L.chunk_idx = queue_pop<kSubs>(q, n, cursor, lane);
            L.fin = kNoFin;
        }
        __syncthreads();
        const uint32_t idx = uniform32(L.chunk_idx);
        if (idx >= n) {
            break;
        }
        const uint64_t chunk = order[idx];
        gcu8* src = global_bytes(srcs[chunk]);
        gu8* dst = global_bytes(dsts[chunk]);
        const uint64_t src_size = uniform64(src_sizes[chunk]);
        const uint64_t cap = uniform64(caps[chunk]);
        if (!plan_two_wave(plan, cost_bin(src_size, cap))) {
            /* the queue holds every two-wave chunk before the one-wave
             * ones, so the block splits for good */
            split_idx = idx;
            break;
        }
        uint64_t sp = 0, declared = 0, dp = 0, flushed = 0, ring_lo = 0;
        uint64_t fuel = src_size == UINT64_MAX ? UINT64_MAX : src_size + 2;
        int status = kOk;
        /* the parser's preamble verdict (Snappy), handed over as a failed
         * window record when it is not OK */
        const int begun = parser ? chunk_begin<kFormat>(src, src_size, cap, sp, declared) : kOk;
        if constexpr (kSlotted) {
            lds_flag* broken = (lds_flag*)&L.broken;
            if (parser) {
                /* each window: wait until its slot's previous record has
                 * run, parse into the slot, publish */
                for (;;) {
                    const uint32_t sl = seq % kS;
                    if (seq >= kS &&
                        !await_flag((const lds_flag*)&L.freed[sl], seq - kS + 1, broken)) {
                        break;
                    }
                    uint32_t touch = 0;
                    const bool last =
                        parse_into<kFormat>(L.pipe.rec[sl], L.stage_of(sl), begun, fuel, src,
                                            src_size, cap, declared, sp, dp, lane, touch);
                    raise_flag((lds_flag*)&L.pub[sl], seq + 1);
                    seq++;
                    /* the look-ahead loads land by now */
                    asm volatile("" ::"v"(touch));
                    if (last) {
                        break;
                    }
                }
            } else {
                /* each window: wait for its record, run it, free the slot */
                for (;;) {
                    const uint32_t sl = seq % kS;
                    if (!await_flag((const lds_flag*)&L.pub[sl], seq + 1, broken)) {
                        status = CUDEC_ERR_CUDA;
                        break;
                    }
                    const bool last =
                        run_record<kParseOnly>(L.pipe.w, L.pipe.rec[sl], L.stage_of(sl), src, dst,
                                               cap, dp, flushed, ring_lo, status, lane);
                    raise_flag((lds_flag*)&L.freed[sl], seq + 1);
                    seq++;
                    if (last) {
                        break;
                    }
                }
            }
        } else {
            bool pdone = false;
            /* step k: the parser fills slot k mod 2 while the copier runs
             * step k - 1's record; one barrier a step */
            for (uint32_t step = 0;; step++) {
                const uint32_t b = step & 1;
                if (parser) {
                    if (!pdone) {
                        uint32_t touch = 0;
                        pdone = parse_into<kFormat>(L.pipe.rec[b], L.stage_of(b), begun, fuel, src,
                                                    src_size, cap, declared, sp, dp, lane, touch);
                        if (pdone) {
                            L.fin = step;
                        }
                        /* the look-ahead loads land by now */
                        asm volatile("" ::"v"(touch));
                    }
                } else if (step > 0 && status == kOk) {
                    run_record<kParseOnly>(L.pipe.w, L.pipe.rec[b ^ 1], L.stage_of(b ^ 1), src,
                                           dst, cap, dp, flushed, ring_lo, status, lane);
                }
                __syncthreads();
                const uint32_t fin = uniform32(L.fin);
                if (fin != kNoFin && step == fin + 1) {
                    break;
                }
            }
        }
        if (!parser) {
            write_result(results, chunk, status, dp);
        }
    }
    if (split_idx == 0xFFFFFFFFu) {
        return;
    }
    /* one wave per chunk from here: the parser decodes the chunk in hand,
     * the other wave takes its own from the queue; no barrier follows */
    WaveLds& w = parser ? L.split.a : L.split.b;
    uint32_t idx = parser ? split_idx : queue_pop<kSubs>(q, n, cursor, lane);
    while (idx < n) {
        const uint64_t chunk = order[idx];
        gcu8* src = global_bytes(srcs[chunk]);
        gu8* dst = global_bytes(dsts[chunk]);
        const uint64_t src_size = uniform64(src_sizes[chunk]);
        const uint64_t cap = uniform64(caps[chunk]);
        const ChunkOut o = decode_chunk<kFormat, kParseOnly>(w, src, src_size, dst, cap, lane);
        write_result(results, chunk, (int)uniform32((uint32_t)o.status), uniform64(o.produced));
        idx = queue_pop<kSubs>(q, n, cursor, lane);
    }
}

/* Launch geometry and a scratch pool, per device, set up on first use. The
 * pool keeps freed scratch (release threshold at its maximum) so that a
 * stream of launches does not return memory to the driver between them. */
struct DeviceState {
    int cus = 0;
    int pipe_blocks_per_cu = 0;  /* piped_batch<_, 4> (two-wave blocks) */
    int split_blocks_per_cu = 0; /* piped_batch<_, 5>: no two-wave chunk, or Snappy */
    Plan plan[2];                /* per format (kFormatLz4, kFormatSnappy) */
    uint64_t min_cost[2];        /* the threshold each plan was last set from */
    hipMemPool_t pool = nullptr; /* made on the first allocating launch */
    bool pool_tried = false;
    void* fixed = nullptr;       /* this device's address of fixed_scratch */
    hipError_t fixed_status = hipErrorNotReady;
    bool ready = false;
};
constexpr int kMaxDevices = 64;

/* Scratch of the allocation-free launch (launch_tuned_fixed): the queue and
 * the order of up to kFixedMaxChunks chunks (4.04 MiB), one copy per device
 * for the life of the process. Host code never passes this name to a
 * kernel or a copy: on the host it is a stand-in, and each device's
 * address of its own copy comes from hipGetSymbolAddress
 * (DeviceState::fixed). */
constexpr size_t kFixedMaxChunks = CUDEC_MI355X_NOALLOC_MAX_CHUNKS;
constexpr size_t kFixedBytes = sizeof(Queue) + sizeof(uint32_t) * kFixedMaxChunks;
static __device__ __align__(256) uint8_t scratch[kFixedBytes];

/* The plan that puts chunks costing at least min_cost on two waves: every
 * bin up to the bin of a chunk whose cost is exactly min_cost (costlier
 * bins have smaller numbers). 0 takes every bin, UINT64_MAX none. */
Plan plan_from_threshold(uint64_t min_cost) {
    Plan p = {};
    const int32_t last = min_cost == 0 ? kBins - 1
                         : min_cost == UINT64_MAX ? -1
                                                  : (int32_t)(kBins - 1 - cost_key(min_cost));
    for (int32_t b = 0; b <= last; b++) {
        p.two_wave[b >> 6] |= 1ull << (b & 63);
    }
    return p;
}
DeviceState g_devices[kMaxDevices];
std::mutex g_devices_lock;

hipError_t device_state(DeviceState** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) {
        return e;
    }
    if (dev < 0 || dev >= kMaxDevices) {
        return hipErrorInvalidDevice;
    }
    std::lock_guard<std::mutex> hold(g_devices_lock);
    DeviceState& d = g_devices[dev];
    if (!d.ready) {
        e = hipDeviceGetAttribute(&d.cus, hipDeviceAttributeMultiprocessorCount, dev);
        if (e != hipSuccess) {
            return e;
        }
        int per = 0;
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, piped_batch<kFormatLz4, 4, 1, false>,
                                                         kPipeBlock, 0);
        if (e != hipSuccess) {
            return e;
        }
        d.pipe_blocks_per_cu = per > 0 ? per : 1;
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, piped_batch<kFormatLz4, 5, 1, false>,
                                                         kPipeBlock, 0);
        if (e != hipSuccess) {
            return e;
        }
        d.split_blocks_per_cu = per > 0 ? per : 1;
        for (int f = 0; f < 2; f++) {
            d.plan[f] = plan_from_threshold(kDefaultPipeMinCost);
            d.min_cost[f] = kDefaultPipeMinCost;
        }
        /* this device is current: its own copy of the fixed scratch */
        d.fixed_status = hipGetSymbolAddress(&d.fixed, HIP_SYMBOL(scratch));
        if (d.fixed_status != hipSuccess) {
            d.fixed = nullptr;
            (void)hipGetLastError(); /* only the allocation-free entry needs it */
        }
        d.ready = true;
    }
    *out = &d;
    return hipSuccess;
}

/* The device's scratch pool, made on first use. It keeps freed scratch up
 * to kPoolKeep bytes, so that a stream of launches does not return memory
 * to the driver between them, while one very large batch's scratch is
 * given back once it is freed. */
constexpr uint64_t kPoolKeep = 64ull << 20;
hipMemPool_t scratch_pool(DeviceState& d, int dev) {
    std::lock_guard<std::mutex> hold(g_devices_lock);
    if (!d.pool_tried) {
        d.pool_tried = true;
        hipMemPoolProps props = {};
        props.allocType = hipMemAllocationTypePinned;
        props.location.type = hipMemLocationTypeDevice;
        props.location.id = dev;
        hipMemPool_t pool = nullptr;
        if (hipMemPoolCreate(&pool, &props) == hipSuccess) {
            uint64_t keep = kPoolKeep;
            if (hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep) == hipSuccess) {
                d.pool = pool;
            } else {
                (void)hipMemPoolDestroy(pool);
            }
        }
        (void)hipGetLastError(); /* a pool that could not be made is not an error */
    }
    return d.pool;
}

}  // namespace

hipError_t set_pipe_min_cost(int format, uint64_t min_cost) {
    DeviceState* dev = nullptr;
    const hipError_t e = device_state(&dev);
    if (e != hipSuccess) {
        return e;
    }
    std::lock_guard<std::mutex> hold(g_devices_lock);
    dev->plan[format] = plan_from_threshold(min_cost);
    dev->min_cost[format] = min_cost;
    return hipSuccess;
}

hipError_t get_pipe_min_cost(int format, uint64_t* min_cost) {
    DeviceState* dev = nullptr;
    const hipError_t e = device_state(&dev);
    if (e != hipSuccess) {
        return e;
    }
    std::lock_guard<std::mutex> hold(g_devices_lock);
    *min_cost = dev->min_cost[format];
    return hipSuccess;
}

hipError_t set_decode_plan(int format, const uint16_t* lanes) {
    Plan p = {};
    for (int b = 0; b < kBins; b++) {
        if (lanes[b] == 128) {
            p.two_wave[b >> 6] |= 1ull << (b & 63);
        }
    }
    DeviceState* dev = nullptr;
    const hipError_t e = device_state(&dev);
    if (e != hipSuccess) {
        return e;
    }
    std::lock_guard<std::mutex> hold(g_devices_lock);
    dev->plan[format] = p;
    return hipSuccess;
}

hipError_t get_decode_plan(int format, uint16_t* lanes) {
    DeviceState* dev = nullptr;
    const hipError_t e = device_state(&dev);
    if (e != hipSuccess) {
        return e;
    }
    Plan p;
    {
        std::lock_guard<std::mutex> hold(g_devices_lock);
        p = dev->plan[format];
    }
    for (int b = 0; b < kBins; b++) {
        lanes[b] = plan_two_wave(p, (uint32_t)b) ? 128 : 64;
    }
    return hipSuccess;
}

uint32_t host_cost_bin(uint64_t src_size, uint64_t cap) {
    const uint64_t lim = 1ull << 61;
    const uint64_t a = src_size < lim ? src_size : lim;
    const uint64_t b = cap < lim ? cap : lim;
    const uint64_t cost = b + (b > a ? b - a : 0);
    return kBins - 1 - cost_key(cost);
}

namespace {

/* Bins the batch into the queue held by `work` (a Queue, then order[n])
 * and launches the persistent decoder, all on `stream`; parse_only
 * launches its parse-only build. */
hipError_t queue_and_decode(DeviceState* dev, int format, const void* const* srcs,
                            const size_t* src_sizes, void* const* dsts, const size_t* caps,
                            size_t n, cudec_chunk_result* results, void* work,
                            hipStream_t stream, bool parse_only) {
    Plan plan;
    {
        std::lock_guard<std::mutex> hold(g_devices_lock);
        plan = dev->plan[format == kFormatSnappy ? kFormatSnappy : kFormatLz4];
    }
    Queue* q = static_cast<Queue*>(work);
    uint32_t* order = reinterpret_cast<uint32_t*>(q + 1);
    const unsigned tb = 256;
    const unsigned nb = (unsigned)((n + tb - 1) / tb);
    hipLaunchKernelGGL(queue_reset, dim3(1), dim3(kKeys), 0, stream, q);
    hipLaunchKernelGGL(queue_count, dim3(nb), dim3(tb), 0, stream, src_sizes, caps, n, q, plan);
    hipLaunchKernelGGL(queue_scan, dim3(1), dim3(1), 0, stream, q);
    hipLaunchKernelGGL(queue_fill, dim3(nb), dim3(tb), 0, stream, src_sizes, caps, n, q, order,
                       plan);
    bool any_two_wave = false;
    for (int w = 0; w < kBins / 64; w++) {
        any_two_wave = any_two_wave || plan.two_wave[w] != 0;
    }
    /* two-wave blocks; a block splits into two one-wave decoders when it
     * meets a one-wave chunk. The five-wave build runs where no chunk is
     * two-wave (every block splits at once) and for Snappy, whose two-wave
     * path fits its 96 VGPRs. */
    const bool five_waves = !any_two_wave || format == kFormatSnappy;
    size_t grid = (size_t)dev->cus *
                  (size_t)(five_waves ? dev->split_blocks_per_cu : dev->pipe_blocks_per_cu);
    if (grid > n) {
        grid = n;
    }
    const dim3 gd((unsigned)grid), bd(kPipeBlock);
    const bool many = (uint64_t)n >= kSubQueueMinPerWave * 2 * (uint64_t)grid;
    const int shape = (parse_only ? 8 : 0) | (format == kFormatSnappy ? 4 : 0) |
                      (five_waves ? 2 : 0) | (many ? 1 : 0);
    switch (shape) {
#define CUDEC_PIPED(kShape, kFmt, kW, kS, kP)                                             \
    case kShape:                                                                          \
        hipLaunchKernelGGL((piped_batch<kFmt, kW, kS, kP>), gd, bd, 0, stream, srcs,      \
                           src_sizes, dsts, caps, n, results, q, order, plan);            \
        break;
        CUDEC_PIPED(0, kFormatLz4, 4, 1, false)
        CUDEC_PIPED(1, kFormatLz4, 4, kSubQueues, false)
        CUDEC_PIPED(2, kFormatLz4, 5, 1, false)
        CUDEC_PIPED(3, kFormatLz4, 5, kSubQueues, false)
        CUDEC_PIPED(6, kFormatSnappy, 5, 1, false)
        CUDEC_PIPED(7, kFormatSnappy, 5, kSubQueues, false)
        CUDEC_PIPED(8, kFormatLz4, 4, 1, true)
        CUDEC_PIPED(9, kFormatLz4, 4, kSubQueues, true)
        CUDEC_PIPED(10, kFormatLz4, 5, 1, true)
        CUDEC_PIPED(11, kFormatLz4, 5, kSubQueues, true)
        CUDEC_PIPED(14, kFormatSnappy, 5, 1, true)
        CUDEC_PIPED(15, kFormatSnappy, 5, kSubQueues, true)
#undef CUDEC_PIPED
        default: /* Snappy always takes the five-wave build */
            return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace

size_t cost_order_bytes(size_t n) {
    return sizeof(Queue) + sizeof(uint32_t) * n;
}

/* The order of a batch by decode cost, costliest first, for a decoder
 * outside this kernel family (the Zstandard frames, whose time also follows
 * the output plus what matches add): the same queue build with every chunk
 * in one class. `work` holds cost_order_bytes(n) bytes; *order points into
 * it (u32 chunk indices, n <= UINT32_MAX). */
hipError_t cost_order(const size_t* sizes, const size_t* caps, size_t n, void* work,
                      hipStream_t stream, const uint32_t** order) {
    if (n > UINT32_MAX) {
        return hipErrorInvalidValue;
    }
    const Plan plan{};
    Queue* q = static_cast<Queue*>(work);
    uint32_t* out = reinterpret_cast<uint32_t*>(q + 1);
    const unsigned tb = 256;
    const unsigned nb = (unsigned)((n + tb - 1) / tb);
    hipLaunchKernelGGL(queue_reset, dim3(1), dim3(kKeys), 0, stream, q);
    hipLaunchKernelGGL(queue_count, dim3(nb), dim3(tb), 0, stream, sizes, caps, n, q, plan);
    hipLaunchKernelGGL(queue_scan, dim3(1), dim3(1), 0, stream, q);
    hipLaunchKernelGGL(queue_fill, dim3(nb), dim3(tb), 0, stream, sizes, caps, n, q, out, plan);
    *order = out;
    return hipGetLastError();
}

hipError_t launch_tuned(int format, const void* const* srcs, const size_t* src_sizes,
                        void* const* dsts, const size_t* caps, size_t n,
                        cudec_chunk_result* results, hipStream_t stream, bool parse_only) {
    if (n > UINT32_MAX) {
        /* the queue indexes chunks in 32 bits; larger batches take the
         * plain form, which has no such limit (and no parse-only build) */
        return parse_only ? hipErrorInvalidValue
                          : launch_plain(format, srcs, src_sizes, dsts, caps, n, results, stream);
    }
    DeviceState* dev = nullptr;
    hipError_t e = device_state(&dev);
    if (e != hipSuccess) {
        return e;
    }
    int id = 0;
    e = hipGetDevice(&id);
    if (e != hipSuccess) {
        return e;
    }
    hipMemPool_t pool = scratch_pool(*dev, id);
    const size_t bytes = sizeof(Queue) + sizeof(uint32_t) * n;
    /* under stream capture the scratch becomes a graph allocation node of
     * the device's default pool; eager launches use the keeping pool */
    hipStreamCaptureStatus capture = hipStreamCaptureStatusNone;
    e = hipStreamIsCapturing(stream, &capture);
    if (e != hipSuccess) {
        return e;
    }
    void* mem = nullptr;
    e = pool != nullptr && capture == hipStreamCaptureStatusNone
            ? hipMallocFromPoolAsync(&mem, bytes, pool, stream)
            : hipMallocAsync(&mem, bytes, stream);
    if (e != hipSuccess) {
        return e;
    }
    e = queue_and_decode(dev, format, srcs, src_sizes, dsts, caps, n, results, mem, stream,
                         parse_only);
    const hipError_t f = hipFreeAsync(mem, stream);
    return e != hipSuccess ? e : f;
}

hipError_t launch_tuned_fixed(int format, const void* const* srcs, const size_t* src_sizes,
                              void* const* dsts, const size_t* caps, size_t n,
                              cudec_chunk_result* results, hipStream_t stream) {
    if (n > kFixedMaxChunks) {
        return hipErrorInvalidValue; /* the entry point refuses this first */
    }
    DeviceState* dev = nullptr;
    hipError_t e = device_state(&dev);
    if (e != hipSuccess) {
        return e;
    }
    if (dev->fixed == nullptr) {
        return dev->fixed_status != hipSuccess ? dev->fixed_status : hipErrorInvalidSymbol;
    }
    return queue_and_decode(dev, format, srcs, src_sizes, dsts, caps, n, results, dev->fixed,
                            stream, false);
}

}  // namespace cudec_amd
