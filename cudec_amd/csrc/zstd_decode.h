/* Zstandard frame decode (RFC 8878), one frame per chunk, written once for
 * the device kernel and for a CPU build of the same code (tests).
 *
 * The frame decode is a serial process: a block's literals section comes
 * first (raw, run-length, or Huffman-coded in one or four streams), then its
 * sequences (literal length, offset value, match length), each decoded from
 * three interleaved FSE states over one backward bitstream and executed in
 * order against the output and the frame's three repeat offsets. Tables
 * survive block boundaries: a Treeless literals section reuses the last
 * Huffman table and a Repeat sequence mode the last FSE table of its field.
 *
 * What runs in parallel is given by the Wave policy (kLanes, lane, sync,
 * uniform: a value every lane holds, as one register for all of them): a
 * wave of 64 lanes on the device, one lane on the host. The serial parts run
 * on every lane with the same values (uniform code); the lane-parallel parts
 * are the four Huffman streams of a literals section, the Huffman table
 * fill, and every byte copy. Lanes exchange results only through memory
 * after a sync(), never through a cross-lane vote, so no convergent
 * operation sits inside the decode's control flow.
 *
 * Accepted frames and verdicts follow the reference's v1 subset
 * (docs/MASTERPLAN.md:1172-1305, tests/zstd_twin_driver.h:178-390): one
 * frame per chunk; no dictionary and a declared content size, else
 * UNSUPPORTED; windows up to 8 MiB; skippable frames UNSUPPORTED; a content
 * size above the capacity, or output past it, OUTPUT_TOO_SMALL; everything
 * else that is refused is CORRUPT_INPUT, including a failed content
 * checksum, bytes after the frame, and bitstreams with bits left over.
 * Within a block every sequence is decoded before the first execution
 * error is reported (tests/zstd_twin_driver.h:268-290), so a bitstream
 * error later in the block takes precedence over it. */
#ifndef CUDEC_AMD_ZSTD_DECODE_H
#define CUDEC_AMD_ZSTD_DECODE_H

#include <stdint.h>

#include "cudec.h"

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define ZD __host__ __device__ __forceinline__
#else
#define ZD inline
#endif

/* Address spaces of the device build: ZG marks global memory (the frames,
 * the output, the literals scratch), ZL the wave's LDS (tables, windows).
 * Named spaces give global_* and ds_* instructions; through generic
 * pointers every access would be a flat one, which waits for outstanding
 * global stores and LDS reads alike. Empty in the CPU build. */
#if defined(__HIP_DEVICE_COMPILE__)
#define ZG __attribute__((address_space(1)))
#define ZL __attribute__((address_space(3)))
#else
#define ZG
#define ZL
#endif

namespace cudec_amd {
namespace zstd {

constexpr uint32_t kMagic = 0xFD2FB528u;
constexpr uint32_t kSkippableMin = 0x184D2A50u;
constexpr uint32_t kSkippableMax = 0x184D2A5Fu;
constexpr uint64_t kWindowMax = 8ull << 20;   /* RFC 8878 3.1.1.1.2 */
constexpr uint32_t kBlockCeiling = 128u << 10; /* RFC 8878 3.1.1.2.4 */
constexpr uint32_t kHufLogMax = 11;            /* literals tables */
constexpr uint32_t kHufCells = 1u << kHufLogMax;
constexpr uint32_t kWeightSymbolMax = 12;      /* Huffman weights alphabet */
constexpr uint32_t kWeightLogMax = 6;
constexpr uint32_t kLitLenMax = 35, kOffsetMax = 31, kMatchLenMax = 52;
constexpr uint32_t kLitLenLog = 9, kOffsetLog = 8, kMatchLenLog = 9;

/* Literal length and match length codes: baseline and extra-bit width
 * (RFC 8878 3.1.1.3.2.1.1). Each baseline is the one before plus 2^width of
 * the one before; a test checks that closed form. */
constexpr uint32_t kLitLenBase[kLitLenMax + 1] = {
    0,  1,  2,  3,  4,  5,  6,  7,  8,  9,  10,  11,  12,  13,   14,   15,   16,   18,
    20, 22, 24, 28, 32, 40, 48, 64, 128, 256, 512, 1024, 2048, 4096, 8192, 16384, 32768, 65536};
constexpr uint8_t kLitLenBits[kLitLenMax + 1] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0,  0,
                                                 0, 0, 0, 0, 1, 1, 1, 1, 2, 2, 3,  3,
                                                 4, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16};
constexpr uint32_t kMatchLenBase[kMatchLenMax + 1] = {
    3,  4,  5,  6,  7,  8,  9,  10, 11, 12,  13,  14,  15,  16,   17,   18,   19,    20,
    21, 22, 23, 24, 25, 26, 27, 28, 29, 30,  31,  32,  33,  34,   35,   37,   39,    41,
    43, 47, 51, 59, 67, 83, 99, 131, 259, 515, 1027, 2051, 4099, 8195, 16387, 32771, 65539};
constexpr uint8_t kMatchLenBits[kMatchLenMax + 1] = {
    0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0,
    0, 0, 0, 0, 0, 1, 1, 1, 1, 2, 2, 3, 3, 4, 4, 5, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16};

/* The predefined distributions (RFC 8878 3.1.1.3.2.2): accuracy logs 6, 5
 * and 6; -1 is a "less than one" probability. */
constexpr int16_t kLitLenDefault[36] = {4, 3, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 2, 1, 1, 1, 2, 2,
                                        2, 2, 2, 2, 2, 2, 2, 3, 2, 1, 1, 1, 1, 1, -1, -1, -1, -1};
constexpr int16_t kOffsetDefault[29] = {1, 1, 1, 1, 1, 1, 2, 2, 2, 1, 1, 1, 1, 1, 1,
                                        1, 1, 1, 1, 1, 1, 1, 1, 1, -1, -1, -1, -1, -1};
constexpr int16_t kMatchLenDefault[53] = {1, 4, 3, 2, 2, 2, 2, 2, 2, 1, 1, 1, 1, 1, 1, 1, 1, 1,
                                          1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1,
                                          1, 1, 1, 1, 1, 1, 1, 1, 1, 1, -1, -1, -1, -1, -1, -1, -1};

ZD uint32_t high_bit(uint32_t v) { return v == 0 ? 0u : 31u - (uint32_t)__builtin_clz(v); }

ZD uint64_t read_le(const ZG uint8_t* p, uint32_t n) {
    uint64_t v = 0;
    for (uint32_t i = 0; i < n; i++) {
        v |= (uint64_t)p[i] << (8 * i);
    }
    return v;
}

/* ------------------------------------------------------------ bitstreams */

/* How a reader's loaded words are held. PerLane: as loaded (a lane decoding
 * its own stream). A reader every lane runs on the same stream takes the
 * Wave policy's Same instead, which may make them one wave-uniform value. */
struct PerLane {
    static ZD uint64_t fix64(uint64_t v) { return v; }
};

/* A backward bitstream (RFC 8878 4.1): read from the last byte towards the
 * first, starting below the highest set bit of the last byte. `acc` holds
 * the next `n` bits in its low bits, the next bit highest; bytes [0, pos)
 * are not loaded yet. */
template <class U>
struct BackBitsT {
    const ZG uint8_t* p;
    uint64_t pos;
    uint64_t acc;
    uint32_t n;
    /* An optional copy of stream bytes [win_lo, win_hi) in faster memory
     * (LDS on the device), which refills read instead of p when it holds
     * the bytes they need. */
    const ZL uint8_t* win;
    uint64_t win_lo, win_hi;

    /* False for an empty stream or a last byte of zero (no start marker). */
    ZD bool start(const ZG uint8_t* s, uint64_t size) {
        p = s;
        win = nullptr;
        win_lo = win_hi = 0;
        if (size == 0 || s[size - 1] == 0) {
            pos = 0;
            acc = 0;
            n = 0;
            return false;
        }
        const uint32_t last = s[size - 1];
        n = high_bit(last);
        acc = last & ((1u << n) - 1u);
        pos = size - 1;
        return true;
    }
    ZD uint64_t remaining() const { return n + 8 * pos; }
    ZD bool at_end() const { return n == 0 && pos == 0; }
    /* Tops the accumulator up to at least 56 bits (or to the stream's
     * start): one 8-byte load takes the next k <= 7 bytes, so a refill costs
     * one memory round trip, not one per byte. */
    ZD void refill() {
        if (n > 56) {
            return;
        }
        if (pos >= 8) {
            uint32_t k = (64 - n) >> 3;
            k = k < 7 ? k : 7;
            uint64_t word = 0; /* bytes pos - 8 .. pos - 1, last on top */
            if (win != nullptr && pos - 8 >= win_lo && pos <= win_hi) {
                const ZL uint8_t* from = win + (pos - 8 - win_lo);
                for (uint32_t j = 0; j < 8; j++) { /* bytes: windows are unaligned */
                    word |= (uint64_t)from[j] << (8 * j);
                }
            } else {
                __builtin_memcpy(&word, (const void*)(p + pos - 8), 8);
            }
            word = U::fix64(word);
            acc = (acc << (8 * k)) | (word >> (64 - 8 * k));
            pos -= k;
            n += 8 * k;
            return;
        }
        while (n <= 56 && pos > 0) {
            pos--;
            const uint64_t b =
                win != nullptr && pos >= win_lo && pos < win_hi ? win[pos - win_lo] : p[pos];
            acc = (acc << 8) | U::fix64(b);
            n += 8;
        }
    }
    /* Whether the next refill would read bytes the window does not hold. */
    ZD bool needs_stage() const {
        if (n > 56 || pos == 0) {
            return false;
        }
        const bool covered = win != nullptr && pos <= win_hi &&
                             (pos >= 8 ? pos - 8 >= win_lo : win_lo == 0);
        return !covered;
    }
    /* Copies stream bytes [pos - cap, pos) (from 0 when pos < cap) into
     * buf and reads from there; one lane, 64 bytes of loads in flight. */
    ZD void stage_lane(ZL uint8_t* buf, uint32_t cap) {
        const uint64_t lo = pos > cap ? pos - cap : 0;
        const uint32_t len = (uint32_t)(pos - lo);
        uint32_t i = 0;
        for (; i + 64 <= len; i += 64) {
            uint8_t v[64];
            __builtin_memcpy(v, (const void*)(p + lo + i), 64);
            for (uint32_t j = 0; j < 64; j++) {
                buf[i + j] = v[j];
            }
        }
        for (; i < len; i++) {
            buf[i] = p[lo + i];
        }
        win = buf;
        win_lo = lo;
        win_hi = pos;
    }
    /* The next c (<= 32) bits; false, reading nothing, if fewer remain. */
    ZD bool read(uint32_t c, uint32_t& v) {
        if (c == 0) {
            v = 0;
            return true;
        }
        if (n < c) {
            refill();
            if (n < c) {
                return false;
            }
        }
        n -= c;
        v = (uint32_t)((acc >> n) & ((1ull << c) - 1ull));
        return true;
    }
    /* The next c bits, zero-filled past the end of the stream. */
    ZD uint32_t peek(uint32_t c) {
        if (n < c) {
            refill();
        }
        const uint64_t mask = (1ull << c) - 1ull;
        return n >= c ? (uint32_t)((acc >> (n - c)) & mask) : (uint32_t)((acc << (c - n)) & mask);
    }
};

using LaneBits = BackBitsT<PerLane>;

/* A forward little-endian bitstream (FSE table descriptions, RFC 8878
 * 4.1.1): bits from the low end of each byte up. */
struct ForwardBits {
    const ZG uint8_t* p;
    uint64_t size;
    uint64_t byte;
    uint32_t bit;

    ZD bool available(uint32_t c) const { return (uint64_t)((bit + c + 7) / 8) <= size - byte; }
    ZD uint32_t peek(uint32_t c) const { /* c <= 16: at most three bytes */
        uint32_t w = 0;
        for (uint32_t k = 0; k < 3; k++) {
            if (byte + k < size) {
                w |= (uint32_t)p[byte + k] << (8 * k);
            }
        }
        return (w >> bit) & ((1u << c) - 1u);
    }
    ZD void skip(uint32_t c) {
        const uint32_t t = bit + c;
        byte += t >> 3;
        bit = t & 7u;
    }
    ZD uint64_t consumed() const { return byte + (bit != 0 ? 1 : 0); }
};

/* ------------------------------------------------------------------- FSE */

struct FseCell {
    uint16_t base; /* next state = base + the bits read */
    uint8_t symbol;
    uint8_t bits;
};

/* Reads an FSE table description (RFC 8878 4.1.1) into counts[0 .. the
 * alphabet's last symbol], -1 for "less than one". Refuses a description
 * that is truncated, whose accuracy log exceeds log_max, that names a
 * symbol past max_symbol, or whose probabilities do not sum to the table
 * size (tests/zstd_fse_twin.cpp pins these rungs in the reference). */
ZD bool read_fse_description(const ZG uint8_t* src, uint64_t size, uint32_t max_symbol,
                             uint32_t log_max, ZL int16_t* counts, uint32_t& last_symbol,
                             uint32_t& log, uint64_t& used) {
    if (size == 0) {
        return false;
    }
    const uint32_t alphabet = max_symbol + 1;
    for (uint32_t s = 0; s < alphabet; s++) {
        counts[s] = 0;
    }
    ForwardBits in{src, size, 0, 0};
    if (!in.available(4)) {
        return false;
    }
    log = in.peek(4) + 5;
    in.skip(4);
    if (log > log_max) {
        return false;
    }
    int32_t remaining = (1 << log) + 1;
    int32_t threshold = 1 << log;
    uint32_t width = log + 1;
    uint32_t symbol = 0;
    bool after_zero = false;
    for (uint32_t round = 0; round < alphabet; round++) {
        if (after_zero) {
            /* runs of zero-probability symbols: 2-bit repeat fields, a
             * field of 3 continuing the run */
            bool open = true;
            for (uint32_t k = 0; k < alphabet && open; k++) {
                if (!in.available(2)) {
                    return false;
                }
                const uint32_t r = in.peek(2);
                in.skip(2);
                symbol += r;
                open = r == 3;
                if (symbol >= alphabet) {
                    return false;
                }
            }
            if (open) {
                return false;
            }
            after_zero = false;
        }
        const int32_t low = 2 * threshold - 1 - remaining; /* values below take width - 1 bits */
        const uint32_t raw = in.peek(width);
        int32_t value = (int32_t)(raw & (uint32_t)(threshold - 1));
        uint32_t take = width - 1;
        if (value >= low) {
            value = (int32_t)(raw & (uint32_t)(2 * threshold - 1));
            if (value >= threshold) {
                value -= low;
            }
            take = width;
        }
        if (!in.available(take)) {
            return false;
        }
        in.skip(take);
        const int32_t probability = value - 1;
        remaining -= probability < 0 ? -probability : probability;
        counts[symbol] = (int16_t)probability;
        symbol++;
        after_zero = probability == 0;
        if (remaining < threshold) {
            if (remaining <= 1) {
                break;
            }
            width = high_bit((uint32_t)remaining) + 1;
            threshold = 1 << (width - 1);
        }
        if (symbol >= alphabet) {
            break;
        }
    }
    if (remaining != 1) {
        return false;
    }
    last_symbol = symbol - 1;
    used = in.consumed();
    return true;
}

/* Builds the decode table of counts[0 .. last_symbol] at accuracy `log`
 * (RFC 8878 4.1.1): "less than one" symbols take the top cells, the rest
 * are spread with step 5/8 of the table + 3, then each cell gets its bit
 * count and next-state base from the running count of its symbol. `next`
 * is scratch of last_symbol + 1 entries. Refuses counts that do not fill
 * the table exactly. */
template <class W>
ZD bool build_fse_table(const W& w, const ZL int16_t* counts, uint32_t last_symbol, uint32_t log,
                        ZL FseCell* cells, ZL uint16_t* next) {
    const uint32_t size = 1u << log;
    uint32_t claimed = 0;
    for (uint32_t s = 0; s <= last_symbol; s++) {
        const int32_t c = counts[s];
        if (c < -1) {
            return false;
        }
        const uint32_t m = c < 0 ? 1u : (uint32_t)c;
        if (m > size - claimed) {
            return false;
        }
        claimed += m;
    }
    if (claimed != size) {
        return false;
    }
    uint32_t high = size - 1;
    for (uint32_t s = 0; s <= last_symbol; s++) {
        if (counts[s] < 0) {
            cells[high].symbol = (uint8_t)s;
            high--;
            next[s] = 1;
        } else {
            next[s] = (uint16_t)counts[s];
        }
    }
    const uint32_t mask = size - 1;
    const uint32_t step = (size >> 1) + (size >> 3) + 3;
    uint32_t at = 0;
    for (uint32_t s = 0; s <= last_symbol; s++) {
        for (int32_t k = 0; k < counts[s]; k++) {
            cells[at].symbol = (uint8_t)s;
            at = (at + step) & mask;
            uint32_t fuel = size;
            while (at > high) {
                if (fuel == 0) {
                    return false;
                }
                fuel--;
                at = (at + step) & mask;
            }
        }
    }
    w.sync();
    for (uint32_t c = 0; c < size; c++) {
        const uint32_t s = cells[c].symbol;
        const uint32_t state = next[s];
        next[s] = (uint16_t)(state + 1);
        const uint32_t bits = log - high_bit(state);
        cells[c].bits = (uint8_t)bits;
        cells[c].base = (uint16_t)((state << bits) - size);
    }
    w.sync();
    return true;
}

/* --------------------------------------------------------------- Huffman */

/* Scratch for a Huffman tree description. */
struct HufScratch {
    int16_t counts[kWeightSymbolMax + 1];
    uint16_t next[kWeightSymbolMax + 1];
    FseCell cells[1u << kWeightLogMax];
    uint8_t weights[256];
};

/* Reads a Huffman tree description (RFC 8878 4.2.1) and fills the decode
 * table: cell index = the next table_log bits; symbol in the low byte, code
 * length in the high byte. Weights come as 4-bit fields (header byte >= 128)
 * or FSE-compressed with two interleaved states (header byte = compressed
 * size); the last symbol's weight is implied by the sum. Refuses weights
 * above 11, a sum that is zero or leaves a non-power-of-two remainder, and
 * fewer than two symbols of weight one. */
template <class W>
ZD bool read_huffman_table(const W& w, const ZG uint8_t* src, uint64_t size, ZL HufScratch& s,
                           ZL uint16_t* table, uint32_t& table_log, uint64_t& used) {
    if (size == 0) {
        return false;
    }
    const uint32_t header = src[0];
    uint32_t written = 0;
    if (header >= 128) {
        written = header - 127;
        const uint64_t payload = (written + 1) / 2;
        if (payload + 1 > size) {
            return false;
        }
        for (uint32_t i = 0; i < written; i++) {
            const uint32_t b = src[1 + i / 2];
            s.weights[i] = (uint8_t)((i & 1) == 0 ? b >> 4 : b & 15u);
        }
        used = payload + 1;
    } else {
        const uint64_t payload = header;
        if (payload + 1 > size) {
            return false;
        }
        uint32_t last = 0, log = 0;
        uint64_t description = 0;
        if (!read_fse_description(src + 1, payload, kWeightSymbolMax, kWeightLogMax, s.counts,
                                  last, log, description)) {
            return false;
        }
        if (!build_fse_table(w, s.counts, last, log, s.cells, s.next)) {
            return false;
        }
        BackBitsT<typename W::Same> in;
        if (!in.start(src + 1 + description, payload - description)) {
            return false;
        }
        const uint32_t cells = 1u << log;
        uint32_t st[2];
        if (!in.read(log, st[0]) || !in.read(log, st[1])) {
            return false;
        }
        /* alternate the two states; when a state's update needs more bits
         * than remain, its symbol and the other state's close the list */
        bool closed = false;
        for (uint32_t k = 0; k < 255; k++) {
            const uint32_t a = k & 1u;
            if (st[a] >= cells) {
                return false;
            }
            const FseCell c = s.cells[st[a]];
            s.weights[written++] = c.symbol;
            if (c.bits > in.remaining()) {
                if (written == 255 || st[a ^ 1u] >= cells) {
                    return false;
                }
                s.weights[written++] = s.cells[st[a ^ 1u]].symbol;
                closed = true;
                break;
            }
            uint32_t v = 0;
            if (!in.read(c.bits, v)) {
                return false;
            }
            st[a] = c.base + v;
        }
        if (!closed) {
            return false;
        }
        used = payload + 1;
    }
    uint32_t total = 0;
    for (uint32_t i = 0; i < written; i++) {
        if (s.weights[i] > kHufLogMax) {
            return false;
        }
        total += (1u << s.weights[i]) >> 1;
    }
    if (total == 0) {
        return false;
    }
    const uint32_t log = high_bit(total) + 1;
    if (log > kHufLogMax) {
        return false;
    }
    const uint32_t rest = (1u << log) - total;
    const uint32_t last_weight = high_bit(rest) + 1;
    if ((1u << (last_weight - 1)) != rest) {
        return false;
    }
    s.weights[written] = (uint8_t)last_weight;
    const uint32_t count = written + 1;
    uint32_t ones = 0;
    uint32_t start[kHufLogMax + 2];
    for (uint32_t k = 0; k <= kHufLogMax + 1; k++) {
        start[k] = 0;
    }
    for (uint32_t i = 0; i < count; i++) {
        ones += s.weights[i] == 1 ? 1u : 0u;
        start[s.weights[i]] += (1u << s.weights[i]) >> 1;
    }
    if (ones < 2) {
        return false;
    }
    /* codes of weight 1 (the longest) take the lowest cells, then weight 2,
     * ...; within a weight, symbols in increasing order */
    uint32_t at = 0;
    for (uint32_t k = 1; k <= log; k++) {
        const uint32_t span = start[k];
        start[k] = at;
        at += span;
    }
    for (uint32_t i = 0; i < count; i++) {
        const uint32_t wt = s.weights[i];
        if (wt == 0) {
            continue;
        }
        const uint32_t span = (1u << wt) >> 1;
        const uint16_t cell = (uint16_t)(i | ((log + 1 - wt) << 8));
        for (uint32_t j = w.lane; j < span; j += W::kLanes) {
            table[start[wt] + j] = cell;
        }
        start[wt] += span;
    }
    w.sync();
    table_log = log;
    return true;
}

/* One Huffman stream of `count` symbols into out: the stream must start
 * with a marker, every code must be paid for by the bits left, and no bit
 * may be left over. */
ZD bool decode_huffman_stream(const ZG uint8_t* src, uint64_t size, const ZL uint16_t* table,
                              uint32_t table_log, ZG uint8_t* out, uint64_t count, ZL uint8_t* win,
                              uint32_t win_cap) {
    LaneBits in;
    if (!in.start(src, size)) {
        return false;
    }
    uint64_t i = 0;
    /* While 8 or more stream bytes are still unloaded, a refill leaves at
     * least 56 bits, which pays for four symbols of at most 11 bits: those
     * decode with no tests, exactly as the loop below would. */
    const uint32_t mask = (1u << table_log) - 1;
    while (i + 4 <= count && in.pos >= 8) {
        if (in.needs_stage()) {
            in.stage_lane(win, win_cap);
        }
        in.refill();
        for (uint32_t j = 0; j < 4; j++) {
            const uint16_t cell = table[(uint32_t)(in.acc >> (in.n - table_log)) & mask];
            in.n -= cell >> 8;
            out[i + j] = (uint8_t)cell;
        }
        i += 4;
    }
    for (; i < count; i++) {
        if (in.needs_stage()) {
            in.stage_lane(win, win_cap); /* the next bytes, into this lane's window */
        }
        const uint16_t cell = table[in.peek(table_log)];
        uint32_t v = 0;
        if (!in.read(cell >> 8, v)) {
            return false;
        }
        out[i] = (uint8_t)cell;
    }
    return in.at_end();
}

/* ------------------------------------------------------------- sequences */

struct SeqTable {
    ZL FseCell* cells;
    uint32_t log;     /* 0 for a run-length table of one cell */
    bool present;
};

/* Up to kBatch decoded and checked sequences of a block are executed
 * together. One sequence's record: distance, literal length, match length,
 * where its literals start in the block's literals, and where its output
 * starts relative to the batch's first output byte. The Wave policy keeps
 * the records (W::Records: put(lane, k, rec), get(k) for every lane,
 * own(lane, k) for the lane that holds record k): on the device lane k
 * holds record k in registers. */
constexpr uint32_t kBatch = 64;
/* A batch's bitstream is at most 64 x (31 + 16 + 16 extra bits + 9 + 9 + 8
 * state bits) = 712 bytes, so a window of this size staged at a batch's
 * start holds every byte its refills read. */
constexpr uint32_t kWindowBytes = 1024;
struct Rec {
    uint64_t d;
    uint32_t ll, ml, lit, at;
};

/* Frame state that survives block boundaries, and per-block scratch. */
struct FrameTables {
    uint16_t huffman[kHufCells];
    FseCell litlen[1u << kLitLenLog];
    FseCell offset[1u << kOffsetLog];
    FseCell matchlen[1u << kMatchLenLog];
    HufScratch huf;
    int16_t counts[kMatchLenMax + 1];
    uint16_t next[kMatchLenMax + 1];
    uint32_t stream_bad[4]; /* per Huffman stream: 1 when it failed */
    /* bitstream windows: a quarter per Huffman stream while literals are
     * decoded, then the whole for the sequences */
    uint8_t window[kWindowBytes];
};

/* Loads one field's table for a sequences section (RFC 8878
 * 3.1.1.3.2.1): 0 predefined, 1 run-length (one symbol byte), 2 an FSE
 * description, 3 repeat the previous table of this field. */
template <class W>
ZD bool load_seq_table(const W& w, uint32_t mode, const ZG uint8_t* src, uint64_t size,
                       uint32_t max_symbol, uint32_t log_max, const int16_t* defaults,
                       uint32_t default_last, uint32_t default_log, SeqTable& t,
                       ZL int16_t* counts, ZL uint16_t* next, uint64_t& used) {
    used = 0;
    if (mode == 3) {
        return t.present;
    }
    if (mode == 1) {
        if (size == 0 || src[0] > max_symbol) {
            return false;
        }
        t.cells[0].base = 0;
        t.cells[0].symbol = src[0];
        t.cells[0].bits = 0;
        w.sync();
        t.log = 0;
        t.present = true;
        used = 1;
        return true;
    }
    uint32_t last = 0, log = 0;
    if (mode == 0) {
        for (uint32_t s = 0; s <= default_last; s++) {
            counts[s] = defaults[s];
        }
        last = default_last;
        log = default_log;
    } else if (!read_fse_description(src, size, max_symbol, log_max, counts, last, log, used)) {
        return false;
    }
    if (!build_fse_table(w, counts, last, log, t.cells, next)) {
        used = 0;
        return false;
    }
    t.log = log;
    t.present = true;
    return true;
}

/* The frame's three repeat offsets (RFC 8878 3.1.1.5). */
struct Repeats {
    uint64_t r0, r1, r2; /* named, not an array: no indexed (memory) access */

    ZD void reset() {
        r0 = 1;
        r1 = 4;
        r2 = 8;
    }
    /* The match distance of an offset value, updating the history; 0 (an
     * error) for a repeat that resolves to zero. Slot 3 stands for a new
     * distance (an explicit offset, or the most recent one minus one). */
    ZD uint64_t resolve(uint64_t value, uint32_t literals) {
        uint64_t d = 0;
        uint32_t slot = 3;
        if (value > 3) {
            d = value - 3;
        } else {
            slot = literals != 0 ? (uint32_t)value - 1 : (uint32_t)value;
            d = slot == 0 ? r0 : slot == 1 ? r1 : slot == 2 ? r2 : (r0 != 0 ? r0 - 1 : 0);
        }
        if (d == 0) {
            return 0;
        }
        if (slot != 0) {
            if (slot != 1) {
                r2 = r1;
            }
            r1 = r0;
            r0 = d;
        }
        return d;
    }
};

/* ----------------------------------------------------------- frame level */

struct FrameHeader {
    uint64_t content_size;
    uint64_t window;
    uint32_t header_size;
    bool checksum;
};

/* RFC 8878 3.1.1.1. CUDEC_OK, CORRUPT_INPUT or UNSUPPORTED. */
ZD int parse_frame_header(const ZG uint8_t* src, uint64_t size, FrameHeader& h) {
    if (size < 4) {
        return CUDEC_ERR_CORRUPT_INPUT;
    }
    const uint32_t magic = (uint32_t)read_le(src, 4);
    if (magic != kMagic) {
        return magic >= kSkippableMin && magic <= kSkippableMax ? CUDEC_ERR_UNSUPPORTED
                                                                 : CUDEC_ERR_CORRUPT_INPUT;
    }
    if (size < 5) {
        return CUDEC_ERR_CORRUPT_INPUT;
    }
    const uint32_t d = src[4];
    const uint32_t fcs_flag = d >> 6;
    const bool single = (d & 0x20u) != 0;
    if ((d & 0x08u) != 0) {
        return CUDEC_ERR_CORRUPT_INPUT; /* reserved bit */
    }
    if ((d & 0x03u) != 0) {
        return CUDEC_ERR_UNSUPPORTED; /* dictionary id */
    }
    const uint32_t fcs_bytes = fcs_flag == 0 ? (single ? 1u : 0u) : 1u << fcs_flag;
    if (fcs_bytes == 0) {
        return CUDEC_ERR_UNSUPPORTED; /* no content size */
    }
    const uint32_t wd_bytes = single ? 0u : 1u;
    h.header_size = 5 + wd_bytes + fcs_bytes;
    if (size < h.header_size) {
        return CUDEC_ERR_CORRUPT_INPUT;
    }
    h.content_size = read_le(src + 5 + wd_bytes, fcs_bytes) + (fcs_bytes == 2 ? 256u : 0u);
    if (single) {
        h.window = h.content_size;
    } else {
        const uint64_t base = 1ull << (10 + (src[5] >> 3));
        h.window = base + (base / 8) * (src[5] & 7u);
        if (h.window > kWindowMax) {
            return CUDEC_ERR_UNSUPPORTED;
        }
    }
    h.checksum = (d & 0x04u) != 0;
    return CUDEC_OK;
}

/* Where a frame decode leaves its content checksum: the caller compares it
 * with the XXH64 of the output (RFC 8878 3.1.1: low 32 bits, seed 0). */
struct FrameEnd {
    uint64_t written;
    uint32_t checksum;
    bool has_checksum;
};

/* Copies n literal bytes to out[0, n): from `lit` (raw or decoded), or the
 * byte `fill` when lit is null (a run-length section). */
template <class W>
ZD void put_literals(const W& w, ZG uint8_t* out, const ZG uint8_t* lit, uint8_t fill,
                     uint64_t n) {
    if (lit != nullptr) {
        for (uint64_t i = w.lane; i < n; i += W::kLanes) {
            out[i] = lit[i];
        }
    } else {
        for (uint64_t i = w.lane; i < n; i += W::kLanes) {
            out[i] = fill;
        }
    }
    w.sync();
}

/* out[at + i] = out[at - d + (i mod d)] for i < n: every source byte lies
 * before `at`, so lanes need no order among themselves. */
template <class W>
ZD void put_match(const W& w, ZG uint8_t* out, uint64_t at, uint64_t d, uint64_t n) {
    const ZG uint8_t* from = out + at - d;
    if (d <= 0xFFFFFFFFull && n <= 0xFFFFFFFFull) {
        const uint32_t d32 = (uint32_t)d;
        for (uint32_t i = w.lane; i < (uint32_t)n; i += W::kLanes) {
            out[at + i] = from[i % d32];
        }
    } else {
        for (uint64_t i = w.lane; i < n; i += W::kLanes) {
            out[at + i] = from[i % d];
        }
    }
    w.sync();
}

/* The wave copies stream bytes [pos - cap, pos) (from 0 when pos < cap)
 * into buf, and the reader reads from there: a sequence batch's bits in one
 * round trip. */
template <class W>
ZD void stage_wave(const W& w, BackBitsT<typename W::Same>& in, ZL uint8_t* buf, uint32_t cap) {
    const uint64_t lo = in.pos > cap ? in.pos - cap : 0;
    const uint32_t len = (uint32_t)(in.pos - lo);
    for (uint32_t i = w.lane * 16; i < len; i += W::kLanes * 16) {
        if (i + 16 <= len) {
            uint8_t v[16];
            __builtin_memcpy(v, (const void*)(in.p + lo + i), 16);
            for (uint32_t j = 0; j < 16; j++) {
                buf[i + j] = v[j];
            }
        } else {
            for (uint32_t j = i; j < len; j++) {
                buf[j] = in.p[lo + j];
            }
        }
    }
    w.sync();
    in.win = buf;
    in.win_lo = lo;
    in.win_hi = in.pos;
}

/* An FSE cell as one 32-bit load, made wave-uniform. */
template <class W>
ZD FseCell load_cell(const W& w, const ZL FseCell* cells, uint32_t state) {
    const FseCell raw = cells[state];
    uint32_t v;
    __builtin_memcpy(&v, &raw, 4);
    v = w.uniform(v);
    FseCell c;
    __builtin_memcpy(&c, &v, 4);
    return c;
}

/* n bytes from src to dst on one lane: 16-byte steps, then single bytes
 * (never past dst + n: the bytes after belong to other lanes). */
ZD void lane_copy(ZG uint8_t* dst, const ZG uint8_t* src, uint32_t n) {
    uint32_t i = 0;
    for (; i + 16 <= n; i += 16) {
        uint8_t v[16];
        __builtin_memcpy(v, (const void*)(src + i), 16);
        __builtin_memcpy((void*)(dst + i), v, 16);
    }
    for (; i < n; i++) {
        dst[i] = src[i];
    }
}

/* Literals and matches a single lane takes alone. */
constexpr uint32_t kLaneBytes = 64;

/* Executes a batch of cnt sequences whose output starts at out[bo]:
 *   1. one lane a sequence: literals of up to kLaneBytes, and matches of up
 *      to kLaneBytes whose source period lies wholly before the batch (so
 *      already written): neither reads anything the batch writes;
 *   2. the wave, sequence by sequence: longer literals, and longer matches
 *      whose period lies before the batch;
 *   3. the wave, in sequence order: every other match, whose source may be
 *      the output of an earlier sequence of the batch (all of which steps 1
 *      and 2 wrote, or step 3 wrote before it).
 * A match's bytes depend only on bytes before its own start, so this order
 * writes what sequence-by-sequence execution writes. */
template <class W>
ZD void execute_batch(const W& w, ZG uint8_t* out, uint64_t bo, const ZG uint8_t* lit,
                      uint8_t fill, const typename W::Records& recs, uint32_t cnt) {
    ZG uint8_t* base = out + bo;
    for (uint32_t k = w.lane; k < cnt; k += W::kLanes) {
        const Rec r = recs.own(w.lane, k);
        if (r.ll <= kLaneBytes) {
            if (lit != nullptr) {
                lane_copy(base + r.at, lit + r.lit, r.ll);
            } else {
                for (uint32_t i = 0; i < r.ll; i++) {
                    base[r.at + i] = fill;
                }
            }
        }
        const uint64_t m = (uint64_t)r.at + r.ll;
        const uint64_t span = r.d < r.ml ? r.d : r.ml;
        if (r.ml <= kLaneBytes && r.d >= m + span) {
            const ZG uint8_t* from = base + m - r.d; /* before the batch: d <= bo + m */
            if (r.d >= r.ml) {
                lane_copy(base + m, from, r.ml);
            } else {
                for (uint32_t i = 0; i < r.ml; i++) {
                    base[m + i] = from[i % (uint32_t)r.d];
                }
            }
        }
    }
    w.sync();
    for (uint32_t k = 0; k < cnt; k++) {
        const Rec r = recs.get(k);
        if (r.ll > kLaneBytes) {
            put_literals(w, base + r.at, lit != nullptr ? lit + r.lit : nullptr, fill, r.ll);
        }
        const uint64_t m = (uint64_t)r.at + r.ll;
        const uint64_t span = r.d < r.ml ? r.d : r.ml;
        if (r.ml > kLaneBytes && r.d >= m + span) {
            put_match(w, out, bo + m, r.d, r.ml);
        }
    }
    for (uint32_t k = 0; k < cnt; k++) {
        const Rec r = recs.get(k);
        const uint64_t m = (uint64_t)r.at + r.ll;
        const uint64_t span = r.d < r.ml ? r.d : r.ml;
        if (r.d < m + span) {
            put_match(w, out, bo + m, r.d, r.ml);
        }
    }
}

/* Decodes one frame of `size` bytes into out[0, cap). `lits` is scratch of
 * kBlockCeiling bytes for Huffman-coded literals. Returns the status; on
 * CUDEC_OK `end` says how much was written and which checksum the output
 * must have. */
template <class W>
ZD int decode_frame(const W& w, const ZG uint8_t* src, uint64_t size, ZG uint8_t* out,
                    uint64_t cap, ZL FrameTables& t, ZG uint8_t* lits, FrameEnd& end) {
    FrameHeader h;
    int st = parse_frame_header(src, size, h);
    if (st != CUDEC_OK) {
        return st;
    }
    if (h.content_size > cap) {
        return CUDEC_ERR_OUTPUT_TOO_SMALL;
    }
    const uint64_t block_max = h.window < kBlockCeiling ? h.window : kBlockCeiling;
    SeqTable ll{t.litlen, 0, false}, of{t.offset, 0, false}, ml{t.matchlen, 0, false};
    uint32_t huf_log = 0; /* 0: no Huffman table yet */
    Repeats rep;
    rep.reset();
    uint64_t pos = h.header_size;
    uint64_t o = 0; /* bytes written */
    for (;;) {
        /* block header (RFC 8878 3.1.1.2) */
        if (size - pos < 3) {
            return CUDEC_ERR_CORRUPT_INPUT;
        }
        const uint32_t bh = (uint32_t)read_le(src + pos, 3);
        const bool last_block = (bh & 1u) != 0;
        const uint32_t type = (bh >> 1) & 3u;
        const uint32_t bsize = bh >> 3;
        if (type == 3 || bsize > block_max) {
            return CUDEC_ERR_CORRUPT_INPUT;
        }
        const uint32_t body_size = type == 1 ? 1u : bsize;
        if (size - pos - 3 < body_size) {
            return CUDEC_ERR_CORRUPT_INPUT;
        }
        const ZG uint8_t* body = src + pos + 3;
        if (type == 0 || type == 1) {
            if (bsize > cap - o) {
                return CUDEC_ERR_OUTPUT_TOO_SMALL;
            }
            put_literals(w, out + o, type == 0 ? body : nullptr, type == 1 ? body[0] : 0, bsize);
            o += bsize;
        } else {
            /* literals section (RFC 8878 3.1.1.3.1) */
            if (body_size == 0) {
                return CUDEC_ERR_CORRUPT_INPUT;
            }
            const uint32_t ltype = body[0] & 3u;
            const uint32_t sf = (body[0] >> 2) & 3u;
            uint32_t lh = 0;
            uint64_t regen = 0, csize = 0;
            uint32_t streams = 0;
            if (ltype <= 1) {
                lh = sf == 1 ? 2u : sf == 3 ? 3u : 1u;
                if (body_size < lh) {
                    return CUDEC_ERR_CORRUPT_INPUT;
                }
                regen = read_le(body, lh) >> (lh == 1 ? 3 : 4);
            } else {
                lh = sf <= 1 ? 3u : sf == 2 ? 4u : 5u;
                const uint32_t fb = sf <= 1 ? 10u : sf == 2 ? 14u : 18u;
                if (body_size < lh) {
                    return CUDEC_ERR_CORRUPT_INPUT;
                }
                const uint64_t raw = read_le(body, lh);
                regen = (raw >> 4) & ((1ull << fb) - 1);
                csize = (raw >> (4 + fb)) & ((1ull << fb) - 1);
                streams = sf == 0 ? 1u : 4u;
            }
            if (regen > block_max) {
                return CUDEC_ERR_CORRUPT_INPUT;
            }
            const uint64_t after = body_size - lh;
            const ZG uint8_t* lit = nullptr; /* null: run-length literals */
            uint8_t fill = 0;
            uint64_t lused = 0;
            if (ltype == 0) {
                if (regen > after) {
                    return CUDEC_ERR_CORRUPT_INPUT;
                }
                lit = body + lh;
                lused = lh + regen;
            } else if (ltype == 1) {
                if (after < 1) {
                    return CUDEC_ERR_CORRUPT_INPUT;
                }
                fill = body[lh];
                lused = lh + 1;
            } else {
                if (csize > after) {
                    return CUDEC_ERR_CORRUPT_INPUT;
                }
                const ZG uint8_t* payload = body + lh;
                uint64_t psize = csize;
                if (ltype == 2) {
                    uint32_t log = 0;
                    uint64_t tree = 0;
                    if (!read_huffman_table(w, payload, psize, t.huf, t.huffman, log, tree)) {
                        return CUDEC_ERR_CORRUPT_INPUT;
                    }
                    huf_log = log;
                    payload += tree;
                    psize -= tree;
                } else if (huf_log == 0) {
                    return CUDEC_ERR_CORRUPT_INPUT; /* treeless with no table */
                }
                if (psize == 0) {
                    return CUDEC_ERR_CORRUPT_INPUT;
                }
                /* one lane a stream; each leaves its verdict in LDS and every
                 * lane reads all four after the sync (no cross-lane vote) */
                for (uint32_t s = w.lane; s < 4; s += W::kLanes) {
                    t.stream_bad[s] = 0;
                }
                w.sync();
                if (streams == 1) {
                    for (uint32_t s = w.lane; s < 1; s += W::kLanes) {
                        t.stream_bad[0] = decode_huffman_stream(payload, psize, t.huffman, huf_log,
                                                                lits, regen, t.window,
                                                                kWindowBytes) ? 0u : 1u;
                    }
                } else {
                    if (psize < 10 || regen < 6) {
                        return CUDEC_ERR_CORRUPT_INPUT;
                    }
                    const uint64_t s0 = read_le(payload, 2), s1 = read_le(payload + 2, 2),
                                   s2 = read_le(payload + 4, 2);
                    const uint64_t bytes = psize - 6;
                    if (s0 + s1 + s2 >= bytes) {
                        return CUDEC_ERR_CORRUPT_INPUT;
                    }
                    const uint64_t seg = (regen + 3) / 4;
                    for (uint32_t s = w.lane; s < 4; s += W::kLanes) {
                        const uint64_t from = 6 + (s > 0 ? s0 : 0) + (s > 1 ? s1 : 0) + (s > 2 ? s2 : 0);
                        const uint64_t len = s == 0 ? s0 : s == 1 ? s1 : s == 2 ? s2 : bytes - s0 - s1 - s2;
                        const uint64_t n = s < 3 ? seg : regen - 3 * seg;
                        t.stream_bad[s] = decode_huffman_stream(
                                              payload + from, len, t.huffman, huf_log,
                                              lits + s * seg, n, t.window + s * (kWindowBytes / 4),
                                              kWindowBytes / 4) ? 0u : 1u;
                    }
                }
                w.sync();
                if ((t.stream_bad[0] | t.stream_bad[1] | t.stream_bad[2] | t.stream_bad[3]) != 0) {
                    return CUDEC_ERR_CORRUPT_INPUT;
                }
                lit = lits;
                lused = lh + csize;
            }

            /* sequences section header (RFC 8878 3.1.1.3.2.1) */
            const ZG uint8_t* sec = body + lused;
            uint64_t left = body_size - lused;
            if (left == 0) {
                return CUDEC_ERR_CORRUPT_INPUT;
            }
            uint32_t count = sec[0];
            uint32_t cb = 1;
            if (count >= 128) {
                if (count < 255) {
                    if (left < 2) {
                        return CUDEC_ERR_CORRUPT_INPUT;
                    }
                    count = ((count - 128) << 8) + sec[1];
                    cb = 2;
                } else {
                    if (left < 3) {
                        return CUDEC_ERR_CORRUPT_INPUT;
                    }
                    count = sec[1] + ((uint32_t)sec[2] << 8) + 0x7F00u;
                    cb = 3;
                }
            }
            uint64_t lp = 0; /* literals consumed */
            if (count != 0) {
                if (count > block_max / 3) {
                    return CUDEC_ERR_CORRUPT_INPUT;
                }
                if (left <= cb) {
                    return CUDEC_ERR_CORRUPT_INPUT;
                }
                const uint32_t modes = sec[cb];
                if ((modes & 3u) != 0) {
                    return CUDEC_ERR_CORRUPT_INPUT;
                }
                sec += cb + 1;
                left -= cb + 1;
                uint64_t u = 0;
                if (!load_seq_table(w, modes >> 6, sec, left, kLitLenMax, kLitLenLog, kLitLenDefault,
                                    35, 6, ll, t.counts, t.next, u)) {
                    return CUDEC_ERR_CORRUPT_INPUT;
                }
                sec += u;
                left -= u;
                if (!load_seq_table(w, (modes >> 4) & 3u, sec, left, kOffsetMax, kOffsetLog,
                                    kOffsetDefault, 28, 5, of, t.counts, t.next, u)) {
                    return CUDEC_ERR_CORRUPT_INPUT;
                }
                sec += u;
                left -= u;
                if (!load_seq_table(w, (modes >> 2) & 3u, sec, left, kMatchLenMax, kMatchLenLog,
                                    kMatchLenDefault, 52, 6, ml, t.counts, t.next, u)) {
                    return CUDEC_ERR_CORRUPT_INPUT;
                }
                sec += u;
                left -= u;

                /* the sequences (RFC 8878 3.1.1.3.2.2 and 4.1) */
                BackBitsT<typename W::Same> in;
                if (left == 0 || !in.start(sec, left)) {
                    return CUDEC_ERR_CORRUPT_INPUT;
                }
                stage_wave(w, in, t.window, kWindowBytes);
                uint32_t sl = 0, so = 0, sm = 0;
                if (!in.read(ll.log, sl) || !in.read(of.log, so) || !in.read(ml.log, sm)) {
                    return CUDEC_ERR_CORRUPT_INPUT;
                }
                const uint32_t nl = 1u << ll.log, no = 1u << of.log, nm = 1u << ml.log;
                int pending = CUDEC_OK; /* the first execution error, reported
                                           once the whole block decodes */
                typename W::Records recs;
                uint32_t cnt = 0;     /* sequences in the batch */
                uint64_t bo = o;      /* the batch's first output byte */
                for (uint32_t k = 0; k < count; k++) {
                    if (k != 0 && (k & (kBatch - 1)) == 0) {
                        stage_wave(w, in, t.window, kWindowBytes);
                    }
                    if (sl >= nl || so >= no || sm >= nm) {
                        return CUDEC_ERR_CORRUPT_INPUT;
                    }
                    const FseCell cl = load_cell(w, ll.cells, sl), co = load_cell(w, of.cells, so),
                                  cm = load_cell(w, ml.cells, sm);
                    if (cl.symbol > kLitLenMax || co.symbol > kOffsetMax || cm.symbol > kMatchLenMax) {
                        return CUDEC_ERR_CORRUPT_INPUT;
                    }
                    uint32_t eo = 0, em = 0, el = 0;
                    if (!in.read(co.symbol, eo) || !in.read(kMatchLenBits[cm.symbol], em) ||
                        !in.read(kLitLenBits[cl.symbol], el)) {
                        return CUDEC_ERR_CORRUPT_INPUT;
                    }
                    const uint64_t value = (1ull << co.symbol) + eo;
                    const uint32_t mlen = kMatchLenBase[cm.symbol] + em;
                    const uint32_t llen = kLitLenBase[cl.symbol] + el;
                    if (k + 1 < count) {
                        uint32_t v = 0;
                        if (!in.read(cl.bits, v)) {
                            return CUDEC_ERR_CORRUPT_INPUT;
                        }
                        sl = cl.base + v;
                        if (!in.read(cm.bits, v)) {
                            return CUDEC_ERR_CORRUPT_INPUT;
                        }
                        sm = cm.base + v;
                        if (!in.read(co.bits, v)) {
                            return CUDEC_ERR_CORRUPT_INPUT;
                        }
                        so = co.base + v;
                    }
                    if (pending != CUDEC_OK) {
                        continue;
                    }
                    /* execute: literals, repeat offset, window, capacity */
                    if (llen > regen - lp) {
                        pending = CUDEC_ERR_CORRUPT_INPUT;
                        continue;
                    }
                    const uint64_t d = rep.resolve(value, llen);
                    const uint64_t m = o + llen;
                    if (d == 0 || d > m || d > h.window) {
                        pending = CUDEC_ERR_CORRUPT_INPUT;
                        continue;
                    }
                    if (m > cap || mlen > cap - m) {
                        pending = CUDEC_ERR_OUTPUT_TOO_SMALL;
                        continue;
                    }
                    recs.put(w.lane, cnt, Rec{d, llen, mlen, (uint32_t)lp, (uint32_t)(o - bo)});
                    cnt++;
                    lp += llen;
                    o = m + mlen;
                    if (cnt == kBatch) {
                        w.sync();
                        execute_batch(w, out, bo, lit, fill, recs, cnt);
                        w.sync();
                        cnt = 0;
                        bo = o;
                    }
                }
                if (!in.at_end()) {
                    return CUDEC_ERR_CORRUPT_INPUT;
                }
                if (pending != CUDEC_OK) {
                    return pending;
                }
                if (cnt != 0) {
                    w.sync();
                    execute_batch(w, out, bo, lit, fill, recs, cnt);
                    w.sync();
                }
            }
            /* the literals after the last sequence */
            if (regen - lp > cap - o) {
                return CUDEC_ERR_OUTPUT_TOO_SMALL;
            }
            put_literals(w, out + o, lit != nullptr ? lit + lp : nullptr, fill, regen - lp);
            o += regen - lp;
        }
        pos += 3 + body_size;
        if (last_block) {
            break;
        }
    }
    if (o != h.content_size) {
        return CUDEC_ERR_CORRUPT_INPUT;
    }
    end.has_checksum = h.checksum;
    end.checksum = 0;
    if (h.checksum) {
        if (size - pos < 4) {
            return CUDEC_ERR_CORRUPT_INPUT;
        }
        end.checksum = (uint32_t)read_le(src + pos, 4);
        pos += 4;
    }
    if (pos != size) {
        return CUDEC_ERR_CORRUPT_INPUT;
    }
    end.written = o;
    return CUDEC_OK;
}

}  // namespace zstd
}  // namespace cudec_amd

#endif
