/* The tuned launch form of the LZ4 batch decoder for gfx950.
 *
 * Work distribution: a persistent grid (LDS-limited residency on every CU)
 * pulls chunk indices from a device queue in which the chunks are binned by
 * compressed size, largest bin first, so the longest decodes start earliest.
 *
 * Per chunk, one wave64 decodes in windows of kWin candidate token positions
 * staged in LDS:
 *   1. stage   -- kStage compressed bytes from the window start are loaded as
 *                 aligned 16-byte blocks (a block holding any byte of the
 *                 chunk never leaves the chunk's pages) into LDS;
 *   2. speculate -- every lane parses 8 candidate positions as if a token
 *                 started there (literal length with one extension byte,
 *                 offset, match length with one extension byte) and records
 *                 where the next token would start;
 *   3. walk    -- a scalar walk follows the true chain from position 0
 *                 through those links (v_readlane), writing each member's
 *                 position into lane j: up to 64 sequences;
 *   4. resolve -- each member lane re-derives its sequence, a wavefront
 *                 prefix sum places its output, and the exact validation
 *                 ladder (src/lz4_block.h:95-258) runs per lane; the first
 *                 member in stream order that rejects, terminates, or would
 *                 overflow the batch ends the batch, so the chunk's status
 *                 is the sequential ladder's;
 *   5. execute -- output goes through a per-wave LDS ring: literals by
 *                 member lanes, independent short matches by member lanes,
 *                 the rest (long runs, matches reading this batch's output)
 *                 in stream order with the whole wave fanned out over the
 *                 bytes (closed-form gather dst[m + i] = dst[m - off +
 *                 (i mod off)]);
 *   6. flush   -- completed ring bytes leave as aligned 16-byte stores; only
 *                 the chunk's first and last partial blocks use byte stores,
 *                 so exactly bytes_written bytes are ever stored.
 * A sequence the window cannot hold (a length with two or more extension
 * bytes, or an output larger than one batch) is decoded by the exact scalar
 * ladder reading global memory, with wide global-to-global copies.
 *
 * Snappy blocks run the plain form's body on the same persistent queue. */
#include <mutex>

#include "common.h"
#include "ladder.h"

namespace cudec_amd {
namespace {

constexpr int kWavesPerBlock = 4;
constexpr int kBlock = kWavesPerBlock * kWave;
constexpr uint32_t kWin = 512;            /* candidate positions per window */
constexpr uint32_t kCand = kWin / kWave;  /* candidates per lane */
constexpr uint32_t kStage = 1024;         /* staged bytes used by a window */
constexpr uint32_t kStageBuf = kStage + 16;
constexpr uint32_t kRing = 4096;          /* per-wave output ring */
constexpr uint32_t kRingMask = kRing - 1;
constexpr uint32_t kBatchMax = 2048;      /* output bytes of one batch */
constexpr uint32_t kStop = 0xFFFFu;       /* link of an unresolvable candidate */
constexpr uint32_t kLaneLit = 64;         /* literal runs copied by their lane */
constexpr uint32_t kLaneMatch = 32;       /* matches copied by their lane (<= 64) */
constexpr int kBins = 64;

static_assert(kBatchMax + 32 <= kRing, "ring must hold a batch and a margin");
static_assert(kLaneMatch <= 64, "a lane match copies at most four 16-byte blocks");
static_assert(kWin + 2 + 269 + 3 <= kStage, "a window's last candidate fits the stage");

struct alignas(16) WaveLds {
    uint8_t ring[kRing];
    uint8_t stage[kStageBuf];
};

/* Scratch of one tuned launch (stream-ordered allocation). */
struct Queue {
    uint32_t bin_count[kBins];
    uint32_t bin_base[kBins];
    uint32_t bin_cursor[kBins];
    uint32_t next;
    uint32_t pad[3];
};

__device__ __forceinline__ uint32_t ring_index(uintptr_t dst_addr, uint64_t pos) {
    return (uint32_t)(dst_addr + pos) & kRingMask;
}

/* ------------------------------------------------------------ staging */

/* Loads compressed bytes [sp, sp + kStage) of the chunk into lds.stage and
 * returns the offset of byte sp inside it. Only aligned 16-byte blocks that
 * start before the chunk end are read (each holds at least one byte of the
 * chunk, so none leaves the chunk's pages); later blocks are zero-filled.
 * Bytes past src_size are never trusted: every member is validated against
 * src_size before anything it read is used. */
__device__ __forceinline__ uint32_t stage_window(WaveLds& w, const uint8_t* src,
                                                 uint64_t src_size, uint64_t sp,
                                                 uint32_t lane) {
    const uintptr_t a = (uintptr_t)(src + sp);
    const uintptr_t base = a & ~(uintptr_t)15;
    const uintptr_t lo = (uintptr_t)src;
    const uintptr_t end = lo + src_size;
    for (uint32_t blk = lane; blk < kStageBuf / 16; blk += kWave) {
        const uintptr_t g = base + 16u * blk;
        uint4 v = make_uint4(0, 0, 0, 0);
        /* only a block that holds at least one byte of [src, src_size) */
        if (g < end && g + 16 > lo) {
            v = *reinterpret_cast<const uint4*>(g);
        }
        *reinterpret_cast<uint4*>(&w.stage[16u * blk]) = v;
    }
    wave_order();
    return (uint32_t)(a - base);
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
 * one extension byte per length. Reads stay below c + 275 < kStage. */
__device__ __forceinline__ Cand parse_candidate(const uint8_t* cw, uint32_t c) {
    Cand k;
    k.tok = cw[c];
    uint32_t p = c + 1;
    uint32_t lit = k.tok >> 4;
    k.complex = false;
    if (lit == 15) {
        const uint32_t e = cw[p++];
        lit += e;
        k.complex = e == 255;
    }
    k.lit = p;
    k.lit_len = lit;
    p += lit;
    k.offset = (uint32_t)cw[p] | ((uint32_t)cw[p + 1] << 8);
    p += 2;
    uint32_t m = (k.tok & 15) + 4;
    if ((k.tok & 15) == 15) {
        const uint32_t e = cw[p++];
        m += e;
        k.complex = k.complex || e == 255;
    }
    k.match_len = m;
    k.next = p;
    return k;
}

/* ------------------------------------------------------------ wave scan */

__device__ __forceinline__ uint32_t wave_inclusive_sum(uint32_t v, uint32_t lane) {
    /* Kogge-Stone over the wave with cross-lane moves; 6 steps. */
#pragma unroll
    for (uint32_t d = 1; d < kWave; d <<= 1) {
        const uint32_t u = (uint32_t)__shfl_up((int)v, d, kWave);
        if (lane >= d) {
            v += u;
        }
    }
    return v;
}

/* ----------------------------------------------------- ring <-> global */

/* Stores ring bytes of positions [from, to) to dst. Full aligned 16-byte
 * blocks go out as one vector store per lane; the partial blocks at either
 * end (only at the chunk's first byte or its last) as single bytes. */
__device__ __forceinline__ void flush_ring(const WaveLds& w, uint8_t* dst,
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
        *reinterpret_cast<uint8_t*>(g) = w.ring[g & kRingMask];
    } else if (lane >= 16 && lane < 16 + tail) {
        const uintptr_t g = b1 + (lane - 16);
        *reinterpret_cast<uint8_t*>(g) = w.ring[g & kRingMask];
    }
    const uint32_t blocks = (uint32_t)((b1 - b0) / 16);
    for (uint32_t b = lane; b < blocks; b += kWave) {
        const uintptr_t g = b0 + 16u * b;
        const uint4 v = *reinterpret_cast<const uint4*>(&w.ring[g & kRingMask]);
        *reinterpret_cast<uint4*>(g) = v;
    }
    wave_order();
}

/* One output byte of history: from the ring when its position is still
 * there, else from global memory (already flushed by this wave). The
 * global read takes the aligned 4-byte word holding the byte: a different
 * load width keeps the compiler from merging the two branches into one
 * flat load, and an aligned word never leaves the page of the byte. */
__device__ __forceinline__ uint8_t history_byte(const WaveLds& w, const uint8_t* dst,
                                                uint64_t y, uint64_t ring_lo) {
    if (y >= ring_lo) {
        return w.ring[ring_index((uintptr_t)dst, y)];
    }
    const uintptr_t a = (uintptr_t)(dst + y);
    const uint32_t word = *reinterpret_cast<const uint32_t*>(a & ~(uintptr_t)3);
    return (uint8_t)(word >> ((a & 3) * 8));
}

/* Byte i (0..15) of a 16-byte register block. */
__device__ __forceinline__ uint32_t block_byte(const uint4& v, uint32_t i) {
    const uint32_t x = (i >> 2) == 0 ? v.x : (i >> 2) == 1 ? v.y : (i >> 2) == 2 ? v.z : v.w;
    return (x >> ((i & 3) * 8)) & 0xFF;
}

/* Exactly n (0..16) bytes of v to ring position r, wrapping at the ring's
 * end: 8/4/2/1-byte stores, so no byte past n is touched. */
__device__ __forceinline__ void ring_put(WaveLds& w, uint32_t r, const uint4& v, uint32_t n) {
    if (n == 0) {
        return;
    }
    if (r + n <= kRing) {
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
        w.ring[(r + i) & kRingMask] = (uint8_t)block_byte(v, i);
    }
}

/* One lane copies a match of len (<= kLaneMatch) bytes at output position
 * m with offset off into the ring; every source byte precedes m or is
 * written earlier by this lane.
 *  - the source period [m - off, m - off + min(len, off)) lies wholly in
 *    flushed history: up to four 16-byte global loads, issued together, so
 *    the whole wave pays one round trip instead of one per byte;
 *  - it lies wholly in the ring with off >= 8: 8-byte ring steps;
 *  - otherwise the period is read bytewise.
 * Bytes past the period repeat it from the ring, 8 bytes a step from a
 * distance d that is a multiple of off and at least 8. */
__device__ __forceinline__ void lane_match_copy(WaveLds& w, const uint8_t* dst, uintptr_t da,
                                                uint64_t m, uint32_t off, uint32_t len,
                                                uint64_t hist_lo, uint64_t cap) {
    const uint64_t y = m - off;
    const uint32_t need = len < off ? len : off;
    const uint32_t rd = ring_index(da, m);
    uint32_t i = 0;
    if (y + need <= hist_lo && y + ((need + 15) & ~15u) <= cap) {
        uint4 h0, h1 = make_uint4(0, 0, 0, 0), h2 = make_uint4(0, 0, 0, 0),
                  h3 = make_uint4(0, 0, 0, 0);
        __builtin_memcpy(&h0, dst + y, 16);
        if (need > 16) {
            __builtin_memcpy(&h1, dst + y + 16, 16);
        }
        if (need > 32) {
            __builtin_memcpy(&h2, dst + y + 32, 16);
        }
        if (need > 48) {
            __builtin_memcpy(&h3, dst + y + 48, 16);
        }
        ring_put(w, rd, h0, need < 16 ? need : 16);
        if (need > 16) {
            ring_put(w, (rd + 16) & kRingMask, h1, need - 16 < 16 ? need - 16 : 16);
        }
        if (need > 32) {
            ring_put(w, (rd + 32) & kRingMask, h2, need - 32 < 16 ? need - 32 : 16);
        }
        if (need > 48) {
            ring_put(w, (rd + 48) & kRingMask, h3, need - 48);
        }
        i = need;
    } else if (y >= hist_lo && rd >= off && rd + len <= kRing && off >= 8) {
        for (; i + 8 <= len; i += 8) {
            uint64_t v;
            __builtin_memcpy(&v, &w.ring[rd - off + i], 8);
            __builtin_memcpy(&w.ring[rd + i], &v, 8);
        }
    } else if (y < hist_lo) {
        for (; i < need; i++) {
            w.ring[(rd + i) & kRingMask] = history_byte(w, dst, y + i, hist_lo);
        }
    }
    if (i < len) {
        /* d for off = 1..7 is 8, 8, 9, 8, 10, 12, 14 */
        const uint32_t d =
            off >= 8 ? off : (uint32_t)((0x0E0C0A08090808ull >> (8 * (off - 1))) & 0xFF);
        for (; i < len && i < d; i++) {
            w.ring[(rd + i) & kRingMask] = w.ring[(rd + i - off) & kRingMask];
        }
        if (rd + len <= kRing) {
            for (; i + 8 <= len; i += 8) {
                uint64_t v;
                __builtin_memcpy(&v, &w.ring[rd + i - d], 8);
                __builtin_memcpy(&w.ring[rd + i], &v, 8);
            }
        }
        for (; i < len; i++) {
            w.ring[(rd + i) & kRingMask] = w.ring[(rd + i - off) & kRingMask];
        }
    }
}

/* --------------------------------------------------- global slow path */

/* dst[0, n) = src[0, n), whole wave, global to global. Destination-aligned
 * 16-byte blocks; each lane's source read is an unaligned 16-byte load that
 * stays inside [src, src + n). */
__device__ __forceinline__ void wave_copy_global(uint8_t* dst, const uint8_t* src,
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
    for (uint64_t b = lane; b < body; b += kWave) {
        uint4 v;
        __builtin_memcpy(&v, src + head + 16 * b, 16);
        *reinterpret_cast<uint4*>(dst + head + 16 * b) = v;
    }
    const uint64_t done = head + 16 * body;
    if (lane < n - done) {
        dst[done + lane] = src[done + lane];
    }
    wave_order();
}

/* dst[m + i] = dst[m - off + (i mod off)] for i < len, whole wave, in
 * global memory; every source byte precedes m and is final. Rounds of at
 * most `off` bytes keep every read behind every write of the same round. */
__device__ __forceinline__ void wave_match_global(uint8_t* dst, uint64_t m, uint64_t off,
                                                  uint64_t len, uint32_t lane) {
    if (off >= 16) {
        /* whole periods: each round copies min(off, 1024) bytes */
        const uint64_t round = off < 1024 ? off : 1024;
        for (uint64_t r0 = 0; r0 < len; r0 += round) {
            const uint64_t n = len - r0 < round ? len - r0 : round;
            wave_copy_global(dst + m + r0, dst + m + r0 - off, n, lane);
        }
        return;
    }
    const uint32_t o = (uint32_t)off;
    const uint64_t from = m - off;
    for (uint64_t i = lane; i < len; i += kWave) {
        dst[m + i] = dst[from + (uint32_t)(i % o)];
    }
    wave_order();
}

/* ------------------------------------------------------------- LZ4 */

struct ChunkOut {
    int status;
    uint64_t produced;
};

__device__ __forceinline__ ChunkOut decode_lz4_chunk(WaveLds& w, const uint8_t* src,
                                                  uint64_t src_size, uint8_t* dst,
                                                  uint64_t cap, uint32_t lane) {
    uint64_t sp = 0, dp = 0;
    uint64_t flushed = 0;  /* output positions below this are in global memory */
    uint64_t ring_lo = 0;  /* output positions at or above this are in the ring */
    int status = kOk;
    bool done = false;
    uint64_t fuel = src_size == UINT64_MAX ? UINT64_MAX : src_size + 2;
    const uintptr_t da = (uintptr_t)dst;

    while (status == kOk && !done) {
        if (fuel-- == 0) {
            status = kCorrupt;
            break;
        }
        const uint32_t delta = stage_window(w, src, src_size, sp, lane);
        const uint8_t* cw = w.stage + delta;

        /* 2. speculate: the link of each candidate position */
        uint32_t link[kCand];
#pragma unroll
        for (uint32_t k = 0; k < kCand; k++) {
            const Cand c = parse_candidate(cw, k * kWave + lane);
            link[k] = c.complex ? kStop : c.next;
        }

        /* 3. walk the true chain */
        uint32_t pos = 0, count = 0;
        uint32_t member = 0; /* lane j: position of member j */
        bool stopped = false;
#pragma unroll
        for (uint32_t k = 0; k < kCand; k++) {
            while (!stopped && pos < (k + 1) * kWave && count < kWave) {
                const uint32_t nx = __builtin_amdgcn_readlane(link[k], pos - k * kWave);
                if (nx == kStop) {
                    stopped = true;
                    break;
                }
                member = lane == count ? pos : member;
                count++;
                pos = nx;
            }
        }
        count = uniform32(count);
        pos = uniform32(pos);

        bool slow = uniform32(count == 0) != 0;
        uint64_t batch_out = 0;
        if (!slow) {
            /* 4. resolve members */
            const bool is_member = lane < count;
            const Cand c = parse_candidate(cw, is_member ? member : 0);
            const uint32_t out_len = is_member ? c.lit_len + c.match_len : 0;
            const uint32_t incl = wave_inclusive_sum(out_len, lane);
            const uint32_t drel = incl - out_len;
            /* the exact ladder for this member's sequence */
            const uint64_t s0 = sp + member; /* absolute source position */
            const uint64_t d0 = dp + drel;   /* absolute destination */
            int ev = kOk;
            bool terminal = false;
            if (is_member) {
                uint64_t s = s0;
                if (s >= src_size) {
                    ev = kCorrupt;
                } else {
                    s++;
                    const uint32_t ln = c.tok >> 4;
                    if (ln == 15) {
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
                        } else if (cap - d0 - L < kLz4SlackDst ||
                                   src_size - s - L < kLz4SlackSrc) {
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
            const bool cut = is_member && drel + mine > kBatchMax;
            const bool flag = is_member && (cut || ev != kOk || terminal);
            const uint64_t flags = __ballot(flag);
            uint32_t take = count; /* members executed in full */
            bool take_terminal = false;
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
            if (status == kOk && !slow) {
                /* 5. execute members [0, take) and the terminal's literal */
                const bool full = lane < take;
                const bool lit_only = take_terminal && lane == take;
                const bool active = full || lit_only;
                const uint32_t last = take_terminal ? take : take - 1;
                const uint32_t end_rel =
                    __builtin_amdgcn_readlane(incl, last) -
                    (take_terminal ? __builtin_amdgcn_readlane(c.match_len, last) : 0);
                batch_out = end_rel;
                const uint64_t D0 = dp;
                const uint64_t D1 = dp + end_rel;
                const uint64_t ring_floor = D1 + 16 > kRing ? D1 + 16 - kRing : 0;
                const uint64_t hist_lo = ring_lo > ring_floor ? ring_lo : ring_floor;
                const uint64_t lit_at = d0;
                const uint64_t m_at = d0 + c.lit_len;
                const bool wide_lit = c.lit_len > kLaneLit;
                /* phase 1: literals, 16 bytes a step while whole steps
                 * fit, then 4 bytes a step; the overrun of at most 3 bytes
                 * lands in this member's own match (>= 4) or, for the
                 * terminal literal, past the batch where no history read
                 * looks */
                if (active && !wide_lit) {
                    uint32_t i0 = 0;
                    for (; i0 + 16 <= c.lit_len; i0 += 16) {
                        uint4 v;
                        __builtin_memcpy(&v, cw + c.lit + i0, 16);
                        const uint32_t r = ring_index(da, lit_at + i0);
                        if (r <= kRing - 16) {
                            __builtin_memcpy(&w.ring[r], &v, 16);
                        } else {
                            for (uint32_t b = 0; b < 16; b++) {
                                w.ring[(r + b) & kRingMask] = (uint8_t)block_byte(v, b);
                            }
                        }
                    }
                    for (uint32_t i = i0; i < c.lit_len; i += 4) {
                        uint32_t v;
                        __builtin_memcpy(&v, cw + c.lit + i, 4);
                        const uint32_t r = ring_index(da, lit_at + i);
                        if (r <= kRing - 4) {
                            __builtin_memcpy(&w.ring[r], &v, 4);
                        } else {
                            for (uint32_t b = 0; b < 4; b++) {
                                w.ring[(r + b) & kRingMask] = (uint8_t)(v >> (8 * b));
                            }
                        }
                    }
                }
                wave_order();
                /* phase 2: matches that read only history or their own
                 * literals, by their own lane */
                const uint64_t y = m_at - c.offset;
                const uint64_t need = c.match_len < c.offset ? c.match_len : c.offset;
                const bool independent = y + need <= D0 || y >= lit_at;
                const bool lane_match = full && !wide_lit && c.match_len <= kLaneMatch &&
                                        independent;
                if (lane_match) {
                    lane_match_copy(w, dst, da, m_at, c.offset, c.match_len, hist_lo, cap);
                }
                wave_order();
                /* phase 3: the rest, in stream order, whole wave */
                uint64_t rest = __ballot(active && (wide_lit || (full && !lane_match)));
                while (rest != 0) {
                    const uint32_t j = __builtin_ctzll(rest);
                    rest &= rest - 1;
                    const uint32_t jl = __builtin_amdgcn_readlane(c.lit_len, j);
                    const uint32_t jlit = __builtin_amdgcn_readlane(c.lit, j);
                    const uint32_t jd = __builtin_amdgcn_readlane(drel, j);
                    const uint64_t jd0 = dp + jd;
                    if (jl > kLaneLit) {
                        for (uint32_t i = lane; i < jl; i += kWave) {
                            w.ring[ring_index(da, jd0 + i)] = cw[jlit + i];
                        }
                    }
                    wave_order();
                    if (j < take) {
                        const uint32_t jm = __builtin_amdgcn_readlane(c.match_len, j);
                        const uint32_t joff = __builtin_amdgcn_readlane(c.offset, j);
                        const bool jdone = __builtin_amdgcn_readlane((uint32_t)lane_match, j) != 0;
                        if (!jdone) {
                            const uint64_t jm0 = jd0 + jl;
                            const uint64_t jfrom = jm0 - joff;
                            for (uint32_t i = lane; i < jm; i += kWave) {
                                const uint64_t yy = jfrom + (i < joff ? i : i % joff);
                                w.ring[ring_index(da, jm0 + i)] = history_byte(w, dst, yy, hist_lo);
                            }
                        }
                    }
                    wave_order();
                }
                /* advance */
                const uint32_t consumed =
                    take_terminal ? (uint32_t)(src_size - sp)
                    : take == count ? pos
                                    : __builtin_amdgcn_readlane(member, take);
                sp = uniform64(sp + consumed);
                dp = uniform64(D1);
                done = take_terminal;
                /* 6. flush */
                /* flush complete 16-byte blocks (everything when done);
                 * computed in address space so an unaligned start cannot
                 * underflow the target */
                const uintptr_t a_end = done ? da + dp : ((da + dp) & ~(uintptr_t)15);
                const uint64_t to = a_end > da + flushed ? (uint64_t)(a_end - da) : flushed;
                if (to > flushed) {
                    flush_ring(w, dst, flushed, to, lane);
                    flushed = to;
                }
                slow = take == 0 && !take_terminal;
            }
        }
        if (status == kOk && !done && slow) {
            /* one sequence by the exact scalar ladder, global memory */
            flush_ring(w, dst, flushed, dp, lane);
            flushed = dp;
            Lz4Cursor cur{src, src_size, cap, sp, dp};
            Lz4Step q;
            bool last = false;
            status = (int)uniform32((uint32_t)lz4_next(cur, q, last));
            last = uniform32(last) != 0;
            if (status == kOk) {
                wave_copy_global(dst + q.lit_dst, src + q.lit_src, q.lit_len, lane);
                if (q.match_len != 0) {
                    wave_match_global(dst, q.match_dst, q.offset, q.match_len, lane);
                }
                sp = uniform64(cur.sp);
                dp = uniform64(cur.dp);
                flushed = dp;
                ring_lo = dp;
                done = last;
            }
        }
        (void)batch_out;
    }
    ChunkOut r;
    r.status = status;
    r.produced = dp;
    return r;
}

/* ------------------------------------------------------------ kernels */

__global__ void queue_reset(Queue* q) {
    const uint32_t t = threadIdx.x;
    if (t < kBins) {
        q->bin_count[t] = 0;
        q->bin_cursor[t] = 0;
    }
    if (t == 0) {
        q->next = 0;
    }
}

/* Bin of a chunk: larger compressed size, smaller bin (decoded first). */
__device__ __forceinline__ uint32_t size_bin(uint64_t size) {
    const uint32_t lg = size == 0 ? 0 : 64u - (uint32_t)__builtin_clzll(size);
    return (kBins - 1) - (lg < kBins - 1 ? lg : kBins - 1);
}

__global__ void queue_count(const size_t* sizes, size_t n, Queue* q) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        atomicAdd(&q->bin_count[size_bin(sizes[i])], 1u);
    }
}

__global__ void queue_scan(Queue* q) {
    if (threadIdx.x == 0) {
        uint32_t acc = 0;
        for (int b = 0; b < kBins; b++) {
            q->bin_base[b] = acc;
            acc += q->bin_count[b];
        }
    }
}

__global__ void queue_fill(const size_t* sizes, size_t n, Queue* q, uint32_t* order) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const uint32_t b = size_bin(sizes[i]);
        order[q->bin_base[b] + atomicAdd(&q->bin_cursor[b], 1u)] = (uint32_t)i;
    }
}

template <int kFormat>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4)))
    tuned_batch(const void* const* srcs, const size_t* src_sizes, void* const* dsts,
                const size_t* caps, size_t n, cudec_chunk_result* results, Queue* q,
                const uint32_t* order) {
    __shared__ WaveLds lds[kWavesPerBlock];
    const uint32_t lane = __lane_id();
    WaveLds& w = lds[threadIdx.x / kWave];
    /* The loop skeleton carries no lane-divergent branch: every lane takes
     * part in the queue atomic (only lane 0 adds 1; the others add 0 and see
     * the same counter), the index is broadcast from the first lane, and
     * every lane stores the same result record. A lane-0-only branch at the
     * loop's top and bottom lets the compiler fold both into a nested
     * exec-masked loop whose exit holds only lane 0. */
    for (;;) {
        const uint32_t mine = atomicAdd(&q->next, lane == 0 ? 1u : 0u);
        const uint32_t idx = uniform32(mine);
        if (idx >= n) {
            break;
        }
        const uint64_t chunk = order[idx];
        const uint8_t* src = uniform_ptr(static_cast<const uint8_t*>(srcs[chunk]));
        uint8_t* dst = uniform_ptr(static_cast<uint8_t*>(dsts[chunk]));
        const uint64_t src_size = uniform64(src_sizes[chunk]);
        const uint64_t cap = uniform64(caps[chunk]);
        int status = kOk;
        uint64_t produced = 0;
        if (kFormat == kFormatLz4) {
            const ChunkOut o = decode_lz4_chunk(w, src, src_size, dst, cap, lane);
            status = o.status;
            produced = o.produced;
        } else {
            SnappyCursor c{src, src_size, cap, 0, 0, 0};
            SnappyStep e;
            bool done = false;
            uint64_t fuel = src_size == UINT64_MAX ? UINT64_MAX : src_size + 1;
            status = snappy_begin(c);
            while (status == kOk) {
                status = snappy_next(c, e, done);
                if (status != kOk) {
                    break;
                }
                if (e.is_copy) {
                    wave_match_global(dst, e.to, e.offset, e.length, lane);
                } else {
                    wave_copy_global(dst + e.to, src + e.from, e.length, lane);
                }
                if (done || fuel-- == 0) {
                    break;
                }
            }
            if (status == kOk && !done) {
                status = kCorrupt;
            }
            produced = c.dp;
        }
        status = (int)uniform32((uint32_t)status);
        write_result(results, chunk, status, uniform64(produced));
    }
}

/* Launch geometry and a scratch pool, per device, set up on first use. The
 * pool keeps freed scratch (release threshold at its maximum) so that a
 * stream of launches does not return memory to the driver between them. */
struct DeviceState {
    int cus = 0;
    int blocks_per_cu = 0;
    hipMemPool_t pool = nullptr;
    bool ready = false;
};
constexpr int kMaxDevices = 64;
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
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, tuned_batch<kFormatLz4>, kBlock, 0);
        if (e != hipSuccess) {
            return e;
        }
        d.blocks_per_cu = per > 0 ? per : 1;
        hipMemPoolProps props = {};
        props.allocType = hipMemAllocationTypePinned;
        props.location.type = hipMemLocationTypeDevice;
        props.location.id = dev;
        hipMemPool_t pool = nullptr;
        if (hipMemPoolCreate(&pool, &props) == hipSuccess) {
            uint64_t keep = UINT64_MAX;
            if (hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep) == hipSuccess) {
                d.pool = pool;
            } else {
                (void)hipMemPoolDestroy(pool);
            }
        }
        (void)hipGetLastError(); /* a pool that could not be made is not an error */
        d.ready = true;
    }
    *out = &d;
    return hipSuccess;
}

}  // namespace

hipError_t launch_tuned(int format, const void* const* srcs, const size_t* src_sizes,
                        void* const* dsts, const size_t* caps, size_t n,
                        cudec_chunk_result* results, hipStream_t stream) {
    if (n > UINT32_MAX) {
        /* the queue indexes chunks in 32 bits; larger batches take the
         * plain form, which has no such limit */
        return launch_plain(format, srcs, src_sizes, dsts, caps, n, results, stream);
    }
    DeviceState* dev = nullptr;
    hipError_t e = device_state(&dev);
    if (e != hipSuccess) {
        return e;
    }
    const size_t bytes = sizeof(Queue) + sizeof(uint32_t) * n;
    void* scratch = nullptr;
    e = dev->pool != nullptr ? hipMallocFromPoolAsync(&scratch, bytes, dev->pool, stream)
                             : hipMallocAsync(&scratch, bytes, stream);
    if (e != hipSuccess) {
        return e;
    }
    Queue* q = static_cast<Queue*>(scratch);
    uint32_t* order = reinterpret_cast<uint32_t*>(q + 1);
    const unsigned tb = 256;
    const unsigned nb = (unsigned)((n + tb - 1) / tb);
    hipLaunchKernelGGL(queue_reset, dim3(1), dim3(kBins), 0, stream, q);
    hipLaunchKernelGGL(queue_count, dim3(nb), dim3(tb), 0, stream, src_sizes, n, q);
    hipLaunchKernelGGL(queue_scan, dim3(1), dim3(1), 0, stream, q);
    hipLaunchKernelGGL(queue_fill, dim3(nb), dim3(tb), 0, stream, src_sizes, n, q, order);
    size_t grid = (size_t)dev->cus * (size_t)dev->blocks_per_cu;
    const size_t need = (n + kWavesPerBlock - 1) / kWavesPerBlock;
    if (grid > need) {
        grid = need;
    }
    if (format == kFormatSnappy) {
        hipLaunchKernelGGL(tuned_batch<kFormatSnappy>, dim3((unsigned)grid), dim3(kBlock), 0,
                           stream, srcs, src_sizes, dsts, caps, n, results, q, order);
    } else {
        hipLaunchKernelGGL(tuned_batch<kFormatLz4>, dim3((unsigned)grid), dim3(kBlock), 0,
                           stream, srcs, src_sizes, dsts, caps, n, results, q, order);
    }
    e = hipGetLastError();
    const hipError_t f = hipFreeAsync(scratch, stream);
    return e != hipSuccess ? e : f;
}

}  // namespace cudec_amd
