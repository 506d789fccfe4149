/* CPU oracle: sequential restatement of iderex/cudec's LZ4 and Snappy block
 * ladders. TEST INFRASTRUCTURE ONLY -- see cudec_oracle.h.
 *
 * Every function names the reference lines it restates. The restatement is
 * pinned (tests/test_oracle_*.py) against the reference's own crafted
 * vectors, fuzz seeds and sweeps, against liblz4 / snappy on corpora and
 * mutants, and against the reference ladders compiled from their own
 * sources (oracle/_ref). */
#define _GNU_SOURCE
#include "cudec_oracle.h"

#include <pthread.h>
#include <sched.h>
#include <stddef.h>
#include <string.h>
#include <time.h>

enum { ST_OK = 0, ST_CORRUPT = 2, ST_TOO_SMALL = 3 };

/* ------------------------------------------------------------------ LZ4 */

/* Constants: src/lz4_block.h:46-60. */
#define LZ4_SLACK_DST 12u    /* MFLIMIT */
#define LZ4_SLACK_SRC 8u     /* 2 offset + 1 token + LASTLITERALS */
#define LZ4_LAST_LITERALS 5u /* LASTLITERALS */
#define LZ4_RUN_MASK 15u
#define LZ4_MIN_MATCH 4u
#define LZ4_LIT_MARGIN 15u   /* liblz4 ilimit margin, literal lengths */
#define LZ4_MATCH_MARGIN 4u  /* liblz4 ilimit margin, match lengths */

typedef struct {
    const uint8_t* src;
    uint64_t src_size;
    uint64_t cap;
    uint64_t sp; /* src_pos */
    uint64_t dp; /* dst_pos */
} lz4_state;

typedef struct {
    uint64_t lit_src, lit_dst, lit_len;
    uint64_t match_src, match_dst, match_len;
} lz4_seq;

/* src/lz4_block.h:95-128 (AccumulateLength): a 255-continued length
 * extension; the capacity bound is checked before each add, the liblz4
 * read margin after each read, and a fuel bound of the remaining source
 * caps the loop. */
static int32_t lz4_accumulate(lz4_state* s, uint64_t* length, uint64_t margin,
                              int initial_check) {
    if (initial_check && s->sp + margin >= s->src_size) {
        return ST_CORRUPT;
    }
    uint64_t fuel = s->src_size - s->sp + 1;
    uint8_t byte;
    do {
        if (fuel-- == 0) {
            return ST_CORRUPT;
        }
        if (s->sp >= s->src_size) {
            return ST_CORRUPT;
        }
        if (*length > s->cap - s->dp) {
            return ST_TOO_SMALL;
        }
        byte = s->src[s->sp++];
        *length += byte;
        if (s->sp + margin > s->src_size) {
            return ST_CORRUPT;
        }
    } while (byte == 255);
    return ST_OK;
}

/* src/lz4_block.h:139-258 (Lz4Parser::Next): token, literal length,
 * literal presence and capacity, the end-of-block terminal rule with the
 * zero-capacity special case, offset (offset == 0 and offset > dst_pos
 * rejected), match length, capacity, and the LASTLITERALS rule. */
static int32_t lz4_next(lz4_state* s, lz4_seq* q, int* done) {
    *done = 0;
    if (s->sp >= s->src_size) {
        return ST_CORRUPT;
    }
    const uint8_t token = s->src[s->sp++];
    uint64_t lit = token >> 4;
    if (lit == LZ4_RUN_MASK) {
        const int32_t st = lz4_accumulate(s, &lit, LZ4_LIT_MARGIN, 1);
        if (st != ST_OK) {
            return st;
        }
    }
    if (lit > s->src_size - s->sp) {
        return ST_CORRUPT;
    }
    if (lit > s->cap - s->dp) {
        return ST_TOO_SMALL;
    }
    const uint64_t dst_slack = s->cap - s->dp - lit;
    const uint64_t src_slack = s->src_size - s->sp - lit;
    if (dst_slack < LZ4_SLACK_DST || src_slack < LZ4_SLACK_SRC) {
        if (s->sp + lit != s->src_size) {
            return ST_CORRUPT;
        }
        if (s->cap == 0 && token != 0) {
            return ST_CORRUPT;
        }
        q->lit_src = s->sp;
        q->lit_dst = s->dp;
        q->lit_len = lit;
        s->sp += lit;
        s->dp += lit;
        q->match_src = 0;
        q->match_dst = s->dp;
        q->match_len = 0;
        *done = 1;
        return ST_OK;
    }
    q->lit_src = s->sp;
    q->lit_dst = s->dp;
    q->lit_len = lit;
    s->sp += lit;
    s->dp += lit;
    const uint64_t offset =
        (uint64_t)s->src[s->sp] | ((uint64_t)s->src[s->sp + 1] << 8);
    s->sp += 2;
    if (offset == 0) {
        return ST_CORRUPT;
    }
    if (offset > s->dp) {
        return ST_CORRUPT;
    }
    uint64_t match = (token & LZ4_RUN_MASK) + LZ4_MIN_MATCH;
    if ((token & LZ4_RUN_MASK) == LZ4_RUN_MASK) {
        const int32_t st = lz4_accumulate(s, &match, LZ4_MATCH_MARGIN, 0);
        if (st != ST_OK) {
            return st;
        }
    }
    if (match > s->cap - s->dp) {
        return ST_TOO_SMALL;
    }
    if (match > s->cap - s->dp - LZ4_LAST_LITERALS) {
        return ST_CORRUPT;
    }
    q->match_src = s->dp - offset;
    q->match_dst = s->dp;
    q->match_len = match;
    s->dp += match;
    return ST_OK;
}

/* Driver: tests/lz4_twin.h:41-77 (sequential execution, chasing copy)
 * with the kernel's fuel cap and result semantics,
 * src/lz4_decode.cuh:90-171 (fuel = src_size + 1, saturating; fuel
 * exhaustion is CORRUPT; bytes_written = dst_pos on OK, else 0). */
int32_t oracle_lz4_decode(const uint8_t* src, uint64_t src_size, uint8_t* dst,
                          uint64_t cap, uint64_t* written) {
    lz4_state s = {src, src_size, cap, 0, 0};
    lz4_seq q;
    int done = 0;
    int32_t st = ST_OK;
    uint64_t fuel = src_size == UINT64_MAX ? UINT64_MAX : src_size + 1;
    for (;;) {
        st = lz4_next(&s, &q, &done);
        if (st != ST_OK) {
            break;
        }
        for (uint64_t i = 0; i < q.lit_len; i++) {
            dst[q.lit_dst + i] = src[q.lit_src + i];
        }
        for (uint64_t i = 0; i < q.match_len; i++) {
            dst[q.match_dst + i] = dst[q.match_src + i];
        }
        if (done || fuel-- == 0) {
            break;
        }
    }
    if (st == ST_OK && !done) {
        st = ST_CORRUPT;
    }
    *written = st == ST_OK ? s.dp : 0;
    return st;
}

/* The same ladder with no copies (tests/parser_twin.cpp:426-440 drives the
 * parser alone at SIZE_MAX capacity), plus the liveness count of
 * tests/termination.cpp:53-107: an OK call must advance src_pos. */
int32_t oracle_lz4_parse(const uint8_t* src, uint64_t src_size, uint64_t cap,
                         uint64_t* dst_pos, uint64_t* sequences,
                         uint64_t* stalls) {
    lz4_state s = {src, src_size, cap, 0, 0};
    lz4_seq q;
    int done = 0;
    int32_t st = ST_OK;
    uint64_t fuel = src_size == UINT64_MAX ? UINT64_MAX : src_size + 1;
    *sequences = 0;
    *stalls = 0;
    for (;;) {
        const uint64_t before = s.sp;
        st = lz4_next(&s, &q, &done);
        if (st != ST_OK) {
            break;
        }
        (*sequences)++;
        if (s.sp == before) {
            (*stalls)++;
        }
        if (done || fuel-- == 0) {
            break;
        }
    }
    if (st == ST_OK && !done) {
        st = ST_CORRUPT;
    }
    *dst_pos = s.dp;
    return st;
}

/* --------------------------------------------------------------- Snappy */

/* Constants: src/snappy_block.h:65-102. */
#define SN_VARINT_GROUPS 5u
#define SN_VARINT_FINAL_MAX 16u
#define SN_LITERAL_INLINE_MAX 59u

typedef struct {
    const uint8_t* src;
    uint64_t src_size;
    uint64_t cap;
    uint64_t sp;
    uint64_t dp;
    uint64_t declared;
    int begun;
    int finished;
    int32_t reject;
} sn_state;

typedef struct {
    uint64_t from, to, length;
    int is_copy;
} sn_elem;

static int32_t sn_refuse(int32_t rung, int32_t status, int32_t* out) {
    if (out) {
        *out = rung;
    }
    return status;
}

/* src/snappy_block.h:190-223 (SnappyDeclaredLength): a varint32 of at most
 * five groups, the fifth refused at or above 16, non-minimal encodings
 * accepted. */
int32_t oracle_snappy_declared(const uint8_t* src, uint64_t src_size,
                               uint64_t* declared, uint64_t* preamble_size,
                               int32_t* reject) {
    uint64_t result = 0;
    for (uint64_t g = 0; g < SN_VARINT_GROUPS; g++) {
        if (g >= src_size) {
            return sn_refuse(ORACLE_SNAPPY_REJECT_PREAMBLE_BYTE_MISSING,
                             ST_CORRUPT, reject);
        }
        const uint8_t b = src[g];
        if (g + 1 == SN_VARINT_GROUPS && b >= SN_VARINT_FINAL_MAX) {
            return sn_refuse(ORACLE_SNAPPY_REJECT_PREAMBLE_FINAL_GROUP,
                             ST_CORRUPT, reject);
        }
        result |= (uint64_t)(b & 0x7F) << (7 * g);
        if ((b & 0x80) == 0) {
            *declared = result;
            *preamble_size = g + 1;
            return ST_OK;
        }
    }
    return sn_refuse(ORACLE_SNAPPY_REJECT_PREAMBLE_UNREACHABLE, ST_CORRUPT,
                     reject);
}

/* src/snappy_block.h:243-263 (SnappyParser::Begin): preamble, then the
 * declared length against the caller's capacity (the only
 * OUTPUT_TOO_SMALL of the ladder). */
static int32_t sn_begin(sn_state* s) {
    if (s->begun) {
        return ST_OK;
    }
    uint64_t pre = 0;
    const int32_t st =
        oracle_snappy_declared(s->src, s->src_size, &s->declared, &pre,
                               &s->reject);
    if (st != ST_OK) {
        return st;
    }
    if (s->declared > s->cap) {
        return sn_refuse(ORACLE_SNAPPY_REJECT_DECLARED_OVER_CAPACITY,
                         ST_TOO_SMALL, &s->reject);
    }
    s->sp = pre;
    s->begun = 1;
    return ST_OK;
}

/* src/snappy_block.h:288-435 (SnappyParser::Next): terminal at exact
 * consumption with exact production; literal (inline length <= 59, else
 * 1-4 little-endian length bytes, length+1 in 64 bits); copy-1/2/4 with
 * offset == 0 and offset > produced refused, and every length bounded by
 * the declared output remaining. */
static int32_t sn_next(sn_state* s, sn_elem* e, int* done) {
    *done = 0;
    if (s->finished) {
        return sn_refuse(ORACLE_SNAPPY_REJECT_PAST_TERMINAL, ST_CORRUPT,
                         &s->reject);
    }
    const int32_t started = sn_begin(s);
    if (started != ST_OK) {
        return started;
    }
    if (s->sp >= s->src_size) {
        if (s->dp != s->declared) {
            return sn_refuse(ORACLE_SNAPPY_REJECT_UNDER_PRODUCTION, ST_CORRUPT,
                             &s->reject);
        }
        e->from = 0;
        e->to = s->dp;
        e->length = 0;
        e->is_copy = 0;
        s->finished = 1;
        *done = 1;
        return ST_OK;
    }
    const uint8_t tag = s->src[s->sp];
    const unsigned kind = tag & 3u;
    if (kind == 0) {
        uint64_t length = tag >> 2;
        uint64_t header = 1;
        if (length > SN_LITERAL_INLINE_MAX) {
            const uint64_t extra = length - SN_LITERAL_INLINE_MAX;
            header += extra;
            if (header > s->src_size - s->sp) {
                return sn_refuse(ORACLE_SNAPPY_REJECT_LITERAL_HEADER_TRUNCATED,
                                 ST_CORRUPT, &s->reject);
            }
            length = 0;
            for (uint64_t i = 0; i < extra; i++) {
                length |= (uint64_t)s->src[s->sp + 1 + i] << (8 * i);
            }
        }
        length += 1;
        s->sp += header;
        if (length > s->src_size - s->sp) {
            return sn_refuse(ORACLE_SNAPPY_REJECT_LITERAL_PAYLOAD_TRUNCATED,
                             ST_CORRUPT, &s->reject);
        }
        if (length > s->declared - s->dp) {
            return sn_refuse(ORACLE_SNAPPY_REJECT_LITERAL_OVER_PRODUCTION,
                             ST_CORRUPT, &s->reject);
        }
        e->from = s->sp;
        e->to = s->dp;
        e->length = length;
        e->is_copy = 0;
        s->sp += length;
        s->dp += length;
        return ST_OK;
    }
    const uint64_t offset_bytes = kind == 1 ? 1 : (kind == 2 ? 2 : 4);
    const uint64_t header = 1 + offset_bytes;
    if (header > s->src_size - s->sp) {
        return sn_refuse(ORACLE_SNAPPY_REJECT_COPY_HEADER_TRUNCATED,
                         ST_CORRUPT, &s->reject);
    }
    uint64_t offset = 0;
    for (uint64_t i = 0; i < offset_bytes; i++) {
        offset |= (uint64_t)s->src[s->sp + 1 + i] << (8 * i);
    }
    uint64_t length;
    if (kind == 1) {
        offset |= (uint64_t)(tag >> 5) << 8;
        length = 4 + ((tag >> 2) & 7u);
    } else {
        length = (uint64_t)(tag >> 2) + 1;
    }
    s->sp += header;
    if (offset == 0) {
        return sn_refuse(ORACLE_SNAPPY_REJECT_COPY_OFFSET_ZERO, ST_CORRUPT,
                         &s->reject);
    }
    if (offset > s->dp) {
        return sn_refuse(ORACLE_SNAPPY_REJECT_COPY_OFFSET_BEFORE_START,
                         ST_CORRUPT, &s->reject);
    }
    if (length > s->declared - s->dp) {
        return sn_refuse(ORACLE_SNAPPY_REJECT_COPY_OVER_PRODUCTION, ST_CORRUPT,
                         &s->reject);
    }
    e->from = s->dp - offset;
    e->to = s->dp;
    e->length = length;
    e->is_copy = 1;
    s->dp += length;
    return ST_OK;
}

/* Driver: tests/snappy_twin.h:140-214 (sequential execution, chasing copy
 * for copies) with the batch contract the reference commits to for its
 * Snappy entry (docs/MASTERPLAN.md:722-733: the LZ4 batch contract, the
 * caller's capacity as the bound) and the LZ4 kernel's fuel and result
 * semantics (src/lz4_decode.cuh:90-171). */
int32_t oracle_snappy_decode(const uint8_t* src, uint64_t src_size,
                             uint8_t* dst, uint64_t cap, uint64_t* written,
                             int32_t* reject) {
    sn_state s;
    memset(&s, 0, sizeof s);
    s.src = src;
    s.src_size = src_size;
    s.cap = cap;
    sn_elem e;
    int done = 0;
    int32_t st = ST_OK;
    uint64_t fuel = src_size == UINT64_MAX ? UINT64_MAX : src_size + 1;
    for (;;) {
        st = sn_next(&s, &e, &done);
        if (st != ST_OK) {
            break;
        }
        if (e.is_copy) {
            for (uint64_t i = 0; i < e.length; i++) {
                dst[e.to + i] = dst[e.from + i];
            }
        } else {
            for (uint64_t i = 0; i < e.length; i++) {
                dst[e.to + i] = src[e.from + i];
            }
        }
        if (done || fuel-- == 0) {
            break;
        }
    }
    if (st == ST_OK && !done) {
        st = ST_CORRUPT;
    }
    if (reject) {
        *reject = st == ST_OK ? ORACLE_SNAPPY_REJECT_NONE : s.reject;
    }
    *written = st == ST_OK ? s.dp : 0;
    return st;
}

int32_t oracle_snappy_parse(const uint8_t* src, uint64_t src_size,
                            uint64_t cap, uint64_t* dst_pos,
                            uint64_t* elements, uint64_t* stalls,
                            int32_t* reject) {
    sn_state s;
    memset(&s, 0, sizeof s);
    s.src = src;
    s.src_size = src_size;
    s.cap = cap;
    sn_elem e;
    int done = 0;
    int32_t st = ST_OK;
    uint64_t fuel = src_size == UINT64_MAX ? UINT64_MAX : src_size + 1;
    *elements = 0;
    *stalls = 0;
    for (;;) {
        const uint64_t before = s.sp;
        const int begun = s.begun;
        st = sn_next(&s, &e, &done);
        if (st != ST_OK) {
            break;
        }
        (*elements)++;
        /* The terminal element consumes nothing by design
         * (src/snappy_block.h:271-282); every other OK call advances. */
        if (!done && begun && s.sp == before) {
            (*stalls)++;
        }
        if (done || fuel-- == 0) {
            break;
        }
    }
    if (st == ST_OK && !done) {
        st = ST_CORRUPT;
    }
    if (st == ST_OK) {
        /* A call past the terminal element must be refused on its own rung
         * (src/snappy_block.h:288-294); anything else is a liveness
         * failure. */
        int again = 0;
        if (sn_next(&s, &e, &again) != ST_CORRUPT ||
            s.reject != ORACLE_SNAPPY_REJECT_PAST_TERMINAL) {
            (*stalls)++;
        }
    }
    if (reject) {
        *reject = st == ST_OK ? ORACLE_SNAPPY_REJECT_NONE : s.reject;
    }
    *dst_pos = s.dp;
    return st;
}

/* ---------------------------------------------------- timed batch decode */

typedef struct {
    int32_t format;
    const uint8_t* const* srcs;
    const uint64_t* src_sizes;
    uint8_t* const* dsts;
    const uint64_t* caps;
    int32_t* status;
    uint64_t* written;
    uint64_t begin, end;
    volatile int* go;
} batch_part;

static void decode_range(batch_part* p) {
    for (uint64_t i = p->begin; i < p->end; i++) {
        if (p->format == 0) {
            p->status[i] = oracle_lz4_decode(p->srcs[i], p->src_sizes[i],
                                             p->dsts[i], p->caps[i],
                                             &p->written[i]);
        } else {
            p->status[i] =
                oracle_snappy_decode(p->srcs[i], p->src_sizes[i], p->dsts[i],
                                     p->caps[i], &p->written[i], NULL);
        }
    }
}

static void* batch_worker(void* arg) {
    batch_part* p = (batch_part*)arg;
    while (!__atomic_load_n(p->go, __ATOMIC_ACQUIRE)) {
        sched_yield();
    }
    decode_range(p);
    return NULL;
}

static double now_seconds(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* Workers are created first and park on a start flag; the clock runs from
 * the flag's release to the last join, so thread creation is outside the
 * timed region. A worker that cannot be created has its range decoded by
 * the calling thread, inside the timed region. */
double oracle_decode_batch_timed(int32_t format, const uint8_t* const* srcs,
                                 const uint64_t* src_sizes,
                                 uint8_t* const* dsts, const uint64_t* caps,
                                 uint64_t n, int32_t threads, int32_t* status,
                                 uint64_t* written) {
    if (threads < 1) {
        threads = 1;
    }
    if (threads > 256) {
        threads = 256;
    }
    if ((uint64_t)threads > n && n > 0) {
        threads = (int32_t)n;
    }
    batch_part parts[256];
    pthread_t tid[256];
    int created[256];
    volatile int go = 0;
    for (int32_t t = 0; t < threads; t++) {
        parts[t].format = format;
        parts[t].srcs = srcs;
        parts[t].src_sizes = src_sizes;
        parts[t].dsts = dsts;
        parts[t].caps = caps;
        parts[t].status = status;
        parts[t].written = written;
        parts[t].begin = n * (uint64_t)t / (uint64_t)threads;
        parts[t].end = n * (uint64_t)(t + 1) / (uint64_t)threads;
        parts[t].go = &go;
        created[t] = 0;
    }
    for (int32_t t = 1; t < threads; t++) {
        created[t] = pthread_create(&tid[t], NULL, batch_worker, &parts[t]) == 0;
    }
    const double t0 = now_seconds();
    __atomic_store_n(&go, 1, __ATOMIC_RELEASE);
    decode_range(&parts[0]);
    for (int32_t t = 1; t < threads; t++) {
        if (created[t]) {
            pthread_join(tid[t], NULL);
        } else {
            decode_range(&parts[t]);
        }
    }
    return now_seconds() - t0;
}
