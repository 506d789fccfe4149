/* Seeded synthetic workloads for the batch decoders, and the CPU compressors
 * that turn them into LZ4 / Snappy block streams.
 *
 * This is workload preparation, not the decode path: the bench and the GPU
 * tests call it to build batches whose every byte is reproducible from a seed.
 * Every sampling step is integer-only (SplitMix64, table-driven geometric,
 * Zipf and log-uniform draws) so the same seed yields the same bytes on any
 * host; no libm call sits on a path that decides a byte.
 *
 * The compressors are liblz4's LZ4_compress_default (acceleration 1),
 * snappy's C API (snappy_compress == snappy::RawCompress) and libzstd's
 * ZSTD_compress (level 3, one frame per chunk), loaded with dlopen so the
 * library builds and loads on a host without them; wl_codec_version
 * reports what was found. */
#define _GNU_SOURCE
#include <dlfcn.h>
#include <pthread.h>
#include <stdatomic.h>
#include <stddef.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

/* ------------------------------------------------------------------ PRNG */

static inline uint64_t splitmix64(uint64_t* state) {
    uint64_t z = (*state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

uint64_t wl_splitmix64_next(uint64_t* state) { return splitmix64(state); }

/* The per-chunk seed: two SplitMix64 rounds over (config seed, index), so
 * neighbouring indices get unrelated streams. */
uint64_t wl_chunk_seed(uint64_t config_seed, uint64_t index) {
    uint64_t s = config_seed ^ (index * 0xD1B54A32D192ED03ull);
    (void)splitmix64(&s);
    return splitmix64(&s);
}

/* Uniform integer in [0, n) by 32x32 multiply-high (n <= 2^32). */
static inline uint32_t below32(uint64_t r, uint32_t n) {
    return (uint32_t)(((r >> 32) * (uint64_t)n) >> 32);
}

/* ------------------------------------------------------------ tables */

/* Zipf(s = 1.1) over 64 symbols, as a 16-bit cumulative table:
 * P(symbol <= k) = kZipf16[k] / 65536, i.e. round(65536 * H_k / H_64) with
 * H_k = sum_{j<=k} j^-1.1. The symbol for a 16-bit uniform u is the first k
 * with u < kZipf16[k]; symbol k is emitted as byte 0x20 + k. */
static const uint16_t kZipf16Minus1[64] = {
    16416, 24074, 28977, 32550, 35345, 37633, 39563, 41230, 42694, 43998,
    45172, 46239, 47217, 48117, 48952, 49730, 50457, 51140, 51784, 52392,
    52969, 53516, 54038, 54536, 55012, 55468, 55905, 56325, 56729, 57119,
    57494, 57857, 58208, 58547, 58876, 59195, 59504, 59804, 60096, 60380,
    60656, 60925, 61187, 61443, 61692, 61935, 62173, 62405, 62632, 62854,
    63071, 63284, 63492, 63696, 63896, 64092, 64285, 64473, 64658, 64840,
    65018, 65194, 65366, 65535};

static uint8_t g_zipf_symbol[65536];

/* 2^(i/64) in Q32, i = 0..64: the sub-octave grid of the log-uniform sizes. */
static const uint64_t kExp2Q32[65] = {
    0x100000000ull, 0x102c9a3e7ull, 0x1059b0d31ull, 0x108745187ull,
    0x10b5586d0ull, 0x10e3ec32dull, 0x111301d01ull, 0x11429aaebull,
    0x1172b83c8ull, 0x11a35beb7ull, 0x11d487317ull, 0x12063b886ull,
    0x12387a6e7ull, 0x126b4565eull, 0x129e9df52ull, 0x12d285a6eull,
    0x1306fe0a3ull, 0x133c08b26ull, 0x1371a7374ull, 0x13a7db34eull,
    0x13dea64c1ull, 0x14160a21full, 0x144e08606ull, 0x1486a2b5cull,
    0x14bfdad53ull, 0x14f9b276aull, 0x15342b56aull, 0x156f4736bull,
    0x15ab07dd5ull, 0x15e76f15bull, 0x16247eb04ull, 0x166238825ull,
    0x16a09e668ull, 0x16dfb23c6ull, 0x171f75e8full, 0x175feb564ull,
    0x17a11473full, 0x17e2f336dull, 0x182589995ull, 0x1868d99b4ull,
    0x18ace5423ull, 0x18f1ae991ull, 0x193737b0dull, 0x197d829feull,
    0x19c49182aull, 0x1a0c667b6ull, 0x1a5503b24ull, 0x1a9e6b558ull,
    0x1ae89f996ull, 0x1b33a2b85ull, 0x1b7f76f30ull, 0x1bcc1e905ull,
    0x1c199bdd8ull, 0x1c67f12e5ull, 0x1cb720dcfull, 0x1d072d4a0ull,
    0x1d5818dd0ull, 0x1da9e603eull, 0x1dfc97338ull, 0x1e502ee79ull,
    0x1ea4afa2aull, 0x1efa1bee6ull, 0x1f50765b7ull, 0x1fa7c181aull,
    0x200000000ull};

/* Geometric on {0, 1, 2, ...} with mean mu: P(G >= k) = (mu / (mu + 1))^k.
 * The survival function is tabulated in Q63 by exact integer recurrence,
 * t[k+1] = floor(t[k] * mu / (mu + 1)); G = max{k : u < t[k]} for a 63-bit
 * uniform u. 4096 entries put the truncated tail below 2^-63 for mu <= 20. */
#define GEO_TABLE 4096
typedef struct {
    uint64_t t[GEO_TABLE];
} geo_table;

static geo_table g_geo12;
static geo_table g_geo20;
static pthread_once_t g_tables_once = PTHREAD_ONCE_INIT;

static void build_geo(geo_table* g, uint64_t mu) {
    unsigned __int128 t = (unsigned __int128)1 << 63;
    for (int k = 0; k < GEO_TABLE; k++) {
        g->t[k] = (uint64_t)t;
        t = t * mu / (mu + 1);
    }
}

static void build_tables(void) {
    unsigned k = 0;
    for (uint32_t u = 0; u < 65536; u++) {
        while (u > kZipf16Minus1[k]) {
            k++;
        }
        g_zipf_symbol[u] = (uint8_t)(0x20 + k);
    }
    build_geo(&g_geo12, 12);
    build_geo(&g_geo20, 20);
}

static inline uint32_t geometric(const geo_table* g, uint64_t r) {
    const uint64_t u = r >> 1;
    /* largest k with u < t[k]; t[0] = 2^63 > u, t is non-increasing */
    uint32_t lo = 0, hi = GEO_TABLE - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if (u < g->t[mid]) {
            lo = mid;
        } else {
            hi = mid - 1;
        }
    }
    return lo;
}

/* --------------------------------------------------------- generators */

/* Chunk kinds. */
enum {
    WL_KIND_TOKENS = 0,     /* the seeded token process, 64-symbol Zipf text */
    WL_KIND_RANDOM = 1,     /* uniform random bytes (literal-only blocks) */
    WL_KIND_REPETITIVE = 2, /* p = 0.05 literals, every offset in [1, 8] */
};

/* Probability thresholds in Q64 (P * 2^64). */
#define Q64(p_num, p_den) ((uint64_t)(((unsigned __int128)(p_num) << 64) / (p_den)))

/* The token process: at each step, with probability p_lit emit a literal run
 * of 1 + Geometric(mean 12) Zipf symbols; otherwise a back-reference copy of
 * 4 + Geometric(mean 20) bytes at an offset uniform in [1, min(pos, max_off)],
 * except that with probability p_short the offset is uniform in
 * [1, min(pos, short_max)]. pos == 0 forces a literal. Copies are executed as
 * a byte-wise chasing copy, so an offset shorter than the length replicates
 * a pattern. Runs and copies are truncated at the chunk end.
 * Draw order per step: one draw for the kind; literal: one draw for the
 * length, then one draw per four symbols (16 bits each); copy: one draw for
 * the length, one for the short/long choice, one for the offset. */
/* What a token-process chunk drew (wl_gen_chunk_stats). */
typedef struct {
    uint64_t steps;          /* steps at pos > 0 (the first step is forced) */
    uint64_t literal_runs;   /* literal steps at pos > 0 */
    uint64_t literal_bytes;  /* bytes of all literal runs, before truncation */
    uint64_t copies;
    uint64_t copy_bytes;     /* before truncation */
    uint64_t short_offsets;  /* copies whose offset is in [1, 16] */
} wl_token_stats;

static void gen_tokens(uint8_t* out, uint64_t size, uint64_t seed,
                       uint64_t p_lit_q64, uint32_t max_off,
                       uint64_t p_short_q64, uint32_t short_max, wl_token_stats* st) {
    uint64_t s = seed;
    uint64_t pos = 0;
    while (pos < size) {
        const uint64_t kind = splitmix64(&s);
        if (st != NULL && pos != 0) {
            st->steps++;
        }
        if (pos == 0 || kind < p_lit_q64) {
            uint64_t len = 1 + geometric(&g_geo12, splitmix64(&s));
            if (st != NULL && pos != 0) {
                st->literal_runs++;
                st->literal_bytes += len;
            }
            if (len > size - pos) {
                len = size - pos;
            }
            uint64_t i = 0;
            while (i < len) {
                uint64_t r = splitmix64(&s);
                for (int q = 0; q < 4 && i < len; q++, i++) {
                    out[pos + i] = g_zipf_symbol[r & 0xFFFF];
                    r >>= 16;
                }
            }
            pos += len;
        } else {
            uint64_t len = 4 + geometric(&g_geo20, splitmix64(&s));
            const uint64_t choice = splitmix64(&s);
            uint64_t window = choice < p_short_q64 ? short_max : max_off;
            if (window > pos) {
                window = pos;
            }
            const uint64_t off = 1 + below32(splitmix64(&s), (uint32_t)window);
            if (st != NULL) {
                st->copies++;
                st->copy_bytes += len;
                st->short_offsets += off <= 16 ? 1 : 0;
            }
            if (len > size - pos) {
                len = size - pos;
            }
            uint8_t* d = out + pos;
            const uint8_t* from = d - off;
            if (off >= len) {
                memcpy(d, from, len);
            } else {
                for (uint64_t i = 0; i < len; i++) {
                    d[i] = from[i];
                }
            }
            pos += len;
        }
    }
}

static void gen_random(uint8_t* out, uint64_t size, uint64_t seed) {
    uint64_t s = seed;
    uint64_t i = 0;
    while (i + 8 <= size) {
        const uint64_t r = splitmix64(&s);
        memcpy(out + i, &r, 8);
        i += 8;
    }
    if (i < size) {
        const uint64_t r = splitmix64(&s);
        memcpy(out + i, &r, (size_t)(size - i));
    }
}

static void gen_chunk(uint8_t* out, uint64_t size, uint64_t seed, int32_t kind,
                      wl_token_stats* st) {
    pthread_once(&g_tables_once, build_tables);
    switch (kind) {
        case WL_KIND_RANDOM:
            gen_random(out, size, seed);
            break;
        case WL_KIND_REPETITIVE:
            gen_tokens(out, size, seed, Q64(5, 100), 8, 0, 8, st);
            break;
        default:
            gen_tokens(out, size, seed, Q64(45, 100), 65535, Q64(30, 100), 16, st);
            break;
    }
}

void wl_gen_chunk(uint8_t* out, uint64_t size, uint64_t seed, int32_t kind) {
    gen_chunk(out, size, seed, kind, NULL);
}

/* wl_gen_chunk, also reporting the draws of a token-process or repetitive
 * chunk into stats[0..5]: steps at pos > 0, literal runs among them,
 * literal bytes, copies, copy bytes, copies at offsets in [1, 16]. The
 * bytes are the same as wl_gen_chunk's. */
void wl_gen_chunk_stats(uint8_t* out, uint64_t size, uint64_t seed, int32_t kind,
                        uint64_t* stats) {
    wl_token_stats st;
    memset(&st, 0, sizeof(st));
    gen_chunk(out, size, seed, kind, &st);
    stats[0] = st.steps;
    stats[1] = st.literal_runs;
    stats[2] = st.literal_bytes;
    stats[3] = st.copies;
    stats[4] = st.copy_bytes;
    stats[5] = st.short_offsets;
}

/* ------------------------------------------------------------- plans */

/* A chunk size log-uniform over [min_size, min_size * 2^octaves): a uniform
 * bucket on a 1/64-octave grid, then a uniform size inside the bucket. */
static uint64_t log_uniform_size(uint64_t* s, uint64_t min_size,
                                 uint32_t octaves) {
    const uint32_t bucket = below32(splitmix64(s), octaves * 64u);
    const uint32_t oct = bucket / 64u, i = bucket % 64u;
    const uint64_t base = min_size << oct;
    const uint64_t lo = (uint64_t)(((unsigned __int128)base * kExp2Q32[i]) >> 32);
    const uint64_t hi = (uint64_t)(((unsigned __int128)base * kExp2Q32[i + 1]) >> 32);
    return lo + (hi > lo ? (splitmix64(s) % (hi - lo)) : 0);
}

uint64_t wl_log_uniform_size(uint64_t* state, uint64_t min_size,
                             uint32_t octaves) {
    return log_uniform_size(state, min_size, octaves);
}

/* The uneven batch: n chunks with sizes log-uniform over
 * [min_size, min_size * 2^octaves); the first pct_random percent of indices
 * are random-byte chunks, the next pct_repetitive percent repetitive, the
 * rest token-process chunks; the order is then shuffled (Fisher-Yates on the
 * same stream). Sizes are drawn before the shuffle, per index, so kinds and
 * sizes are independent. */
void wl_plan_uneven(uint64_t n, uint64_t seed, uint64_t min_size,
                    uint32_t octaves, uint32_t pct_random,
                    uint32_t pct_repetitive, uint64_t* sizes, int32_t* kinds,
                    uint64_t* seeds) {
    uint64_t s = seed;
    const uint64_t n_random = n * pct_random / 100;
    const uint64_t n_rep = n * pct_repetitive / 100;
    for (uint64_t i = 0; i < n; i++) {
        sizes[i] = log_uniform_size(&s, min_size, octaves);
        kinds[i] = i < n_random ? WL_KIND_RANDOM
                   : i < n_random + n_rep ? WL_KIND_REPETITIVE
                                          : WL_KIND_TOKENS;
        seeds[i] = wl_chunk_seed(seed, i);
    }
    for (uint64_t i = n; i > 1; i--) {
        const uint64_t j = splitmix64(&s) % i;
        const uint64_t ts = sizes[i - 1];
        sizes[i - 1] = sizes[j];
        sizes[j] = ts;
        const int32_t tk = kinds[i - 1];
        kinds[i - 1] = kinds[j];
        kinds[j] = tk;
        const uint64_t tz = seeds[i - 1];
        seeds[i - 1] = seeds[j];
        seeds[j] = tz;
    }
}

/* ------------------------------------------------------- thread pool */

typedef struct {
    _Atomic uint64_t next;
    uint64_t n;
    void (*fn)(void* ctx, uint64_t i);
    void* ctx;
} pool_job;

static void* pool_worker(void* arg) {
    pool_job* job = (pool_job*)arg;
    for (;;) {
        const uint64_t i = atomic_fetch_add(&job->next, 1);
        if (i >= job->n) {
            break;
        }
        job->fn(job->ctx, i);
    }
    return NULL;
}

static void run_pool(uint64_t n, int threads, void (*fn)(void*, uint64_t),
                     void* ctx) {
    pool_job job;
    atomic_init(&job.next, 0);
    job.n = n;
    job.fn = fn;
    job.ctx = ctx;
    if (threads < 1) {
        threads = 1;
    }
    if (threads > 256) {
        threads = 256;
    }
    pthread_t tid[256];
    int started = 0;
    for (int t = 1; t < threads; t++) {
        if (pthread_create(&tid[started], NULL, pool_worker, &job) == 0) {
            started++;
        }
    }
    pool_worker(&job);
    for (int t = 0; t < started; t++) {
        pthread_join(tid[t], NULL);
    }
}

typedef struct {
    uint8_t* arena;
    const uint64_t* offsets;
    const uint64_t* sizes;
    const uint64_t* seeds;
    const int32_t* kinds;
} gen_ctx;

static void gen_one(void* p, uint64_t i) {
    gen_ctx* c = (gen_ctx*)p;
    wl_gen_chunk(c->arena + c->offsets[i], c->sizes[i], c->seeds[i],
                 c->kinds[i]);
}

void wl_gen_batch(uint8_t* arena, const uint64_t* offsets,
                  const uint64_t* sizes, const uint64_t* seeds,
                  const int32_t* kinds, uint64_t n, int32_t threads) {
    pthread_once(&g_tables_once, build_tables);
    gen_ctx c = {arena, offsets, sizes, seeds, kinds};
    run_pool(n, threads, gen_one, &c);
}

/* ------------------------------------------------------- compressors */

enum { WL_CODEC_LZ4 = 0, WL_CODEC_SNAPPY = 1, WL_CODEC_ZSTD = 2 };
#define WL_ZSTD_LEVEL 3

typedef int (*lz4_compress_fn)(const char*, char*, int, int);
typedef int (*lz4_decompress_fn)(const char*, char*, int, int);
typedef int (*lz4_bound_fn)(int);
typedef const char* (*lz4_version_fn)(void);
typedef int (*snappy_compress_fn)(const char*, size_t, char*, size_t*);
typedef int (*snappy_uncompress_fn)(const char*, size_t, char*, size_t*);
typedef size_t (*snappy_max_fn)(size_t);
typedef size_t (*zstd_compress_fn)(void*, size_t, const void*, size_t, int);
typedef size_t (*zstd_decompress_fn)(void*, size_t, const void*, size_t);
typedef size_t (*zstd_bound_fn)(size_t);
typedef unsigned (*zstd_is_error_fn)(size_t);
typedef const char* (*zstd_version_fn)(void);

static struct {
    void* lz4;
    void* snappy;
    lz4_compress_fn lz4_compress;
    lz4_decompress_fn lz4_decompress;
    lz4_bound_fn lz4_bound;
    lz4_version_fn lz4_version;
    snappy_compress_fn snappy_compress;
    snappy_uncompress_fn snappy_uncompress;
    snappy_max_fn snappy_max;
    char snappy_path[512];
    void* zstd;
    zstd_compress_fn zstd_compress;
    zstd_decompress_fn zstd_decompress;
    zstd_bound_fn zstd_bound;
    zstd_is_error_fn zstd_is_error;
    zstd_version_fn zstd_version;
} g_codecs;
static pthread_once_t g_codecs_once = PTHREAD_ONCE_INIT;

static void* open_first(const char* const* names) {
    for (int i = 0; names[i]; i++) {
        void* h = dlopen(names[i], RTLD_NOW | RTLD_LOCAL);
        if (h) {
            return h;
        }
    }
    return NULL;
}

static void load_codecs(void) {
    static const char* const lz4_names[] = {
        "liblz4.so.1", "/opt/conda/lib/liblz4.so.1",
        "/usr/lib/x86_64-linux-gnu/liblz4.so.1", NULL};
    static const char* const snappy_names[] = {
        "libsnappy.so.1", "/opt/conda/lib/libsnappy.so.1",
        "/usr/lib/x86_64-linux-gnu/libsnappy.so.1", NULL};
    g_codecs.lz4 = open_first(lz4_names);
    if (g_codecs.lz4) {
        g_codecs.lz4_compress =
            (lz4_compress_fn)dlsym(g_codecs.lz4, "LZ4_compress_default");
        g_codecs.lz4_decompress =
            (lz4_decompress_fn)dlsym(g_codecs.lz4, "LZ4_decompress_safe");
        g_codecs.lz4_bound = (lz4_bound_fn)dlsym(g_codecs.lz4, "LZ4_compressBound");
        g_codecs.lz4_version =
            (lz4_version_fn)dlsym(g_codecs.lz4, "LZ4_versionString");
    }
    g_codecs.snappy = open_first(snappy_names);
    if (g_codecs.snappy) {
        g_codecs.snappy_compress =
            (snappy_compress_fn)dlsym(g_codecs.snappy, "snappy_compress");
        g_codecs.snappy_uncompress =
            (snappy_uncompress_fn)dlsym(g_codecs.snappy, "snappy_uncompress");
        g_codecs.snappy_max =
            (snappy_max_fn)dlsym(g_codecs.snappy, "snappy_max_compressed_length");
        Dl_info info;
        if (g_codecs.snappy_compress &&
            dladdr((void*)g_codecs.snappy_compress, &info) && info.dli_fname) {
            /* the file the soname link names carries the version */
            char* real = realpath(info.dli_fname, NULL);
            snprintf(g_codecs.snappy_path, sizeof g_codecs.snappy_path, "%s",
                     real ? real : info.dli_fname);
            free(real);
        }
    }
    static const char* const zstd_names[] = {
        "/opt/conda/lib/libzstd.so.1", "libzstd.so.1",
        "/usr/lib/x86_64-linux-gnu/libzstd.so.1", NULL};
    g_codecs.zstd = open_first(zstd_names);
    if (g_codecs.zstd) {
        g_codecs.zstd_compress = (zstd_compress_fn)dlsym(g_codecs.zstd, "ZSTD_compress");
        g_codecs.zstd_decompress = (zstd_decompress_fn)dlsym(g_codecs.zstd, "ZSTD_decompress");
        g_codecs.zstd_bound = (zstd_bound_fn)dlsym(g_codecs.zstd, "ZSTD_compressBound");
        g_codecs.zstd_is_error = (zstd_is_error_fn)dlsym(g_codecs.zstd, "ZSTD_isError");
        g_codecs.zstd_version = (zstd_version_fn)dlsym(g_codecs.zstd, "ZSTD_versionString");
    }
}

int32_t wl_codec_available(int32_t codec) {
    pthread_once(&g_codecs_once, load_codecs);
    if (codec == WL_CODEC_LZ4) {
        return g_codecs.lz4_compress && g_codecs.lz4_decompress &&
               g_codecs.lz4_bound;
    }
    if (codec == WL_CODEC_SNAPPY) {
        return g_codecs.snappy_compress && g_codecs.snappy_uncompress &&
               g_codecs.snappy_max;
    }
    if (codec == WL_CODEC_ZSTD) {
        return g_codecs.zstd_compress && g_codecs.zstd_decompress && g_codecs.zstd_bound &&
               g_codecs.zstd_is_error;
    }
    return 0;
}

/* A human-readable name of the library found ("liblz4 1.9.3",
 * "libsnappy (/opt/conda/lib/libsnappy.so.1)", "libzstd 1.4.9 level 3") or
 * "absent". */
const char* wl_codec_version(int32_t codec) {
    static char buf[3][600];
    pthread_once(&g_codecs_once, load_codecs);
    if (codec == WL_CODEC_LZ4) {
        if (!g_codecs.lz4_compress) {
            return "absent";
        }
        snprintf(buf[0], sizeof buf[0], "liblz4 %s",
                 g_codecs.lz4_version ? g_codecs.lz4_version() : "?");
        return buf[0];
    }
    if (codec == WL_CODEC_ZSTD) {
        if (!g_codecs.zstd_compress) {
            return "absent";
        }
        snprintf(buf[2], sizeof buf[2], "libzstd %s level %d",
                 g_codecs.zstd_version ? g_codecs.zstd_version() : "?", WL_ZSTD_LEVEL);
        return buf[2];
    }
    if (!g_codecs.snappy_compress) {
        return "absent";
    }
    snprintf(buf[1], sizeof buf[1], "libsnappy (%s)", g_codecs.snappy_path);
    return buf[1];
}

uint64_t wl_compress_bound(int32_t codec, uint64_t n) {
    pthread_once(&g_codecs_once, load_codecs);
    if (codec == WL_CODEC_LZ4 && g_codecs.lz4_bound && n <= 0x7E000000ull) {
        return (uint64_t)g_codecs.lz4_bound((int)n);
    }
    if (codec == WL_CODEC_SNAPPY && g_codecs.snappy_max) {
        return (uint64_t)g_codecs.snappy_max((size_t)n);
    }
    if (codec == WL_CODEC_ZSTD && g_codecs.zstd_bound) {
        return (uint64_t)g_codecs.zstd_bound((size_t)n);
    }
    return 0;
}

/* One-chunk compression; returns the compressed size, 0 on failure. */
uint64_t wl_compress(int32_t codec, const uint8_t* src, uint64_t n,
                     uint8_t* dst, uint64_t cap) {
    pthread_once(&g_codecs_once, load_codecs);
    if (codec == WL_CODEC_LZ4) {
        if (!g_codecs.lz4_compress || n > 0x7E000000ull || cap > 0x7FFFFFFFull) {
            return 0;
        }
        const int w = g_codecs.lz4_compress((const char*)src, (char*)dst,
                                            (int)n, (int)cap);
        return w > 0 ? (uint64_t)w : 0;
    }
    if (codec == WL_CODEC_SNAPPY) {
        if (!g_codecs.snappy_compress || cap < g_codecs.snappy_max((size_t)n)) {
            return 0;
        }
        size_t out = (size_t)cap;
        return g_codecs.snappy_compress((const char*)src, (size_t)n,
                                        (char*)dst, &out) == 0
                   ? (uint64_t)out
                   : 0;
    }
    if (codec == WL_CODEC_ZSTD && g_codecs.zstd_compress) {
        const size_t w = g_codecs.zstd_compress(dst, (size_t)cap, src, (size_t)n, WL_ZSTD_LEVEL);
        return g_codecs.zstd_is_error(w) ? 0 : (uint64_t)w;
    }
    return 0;
}

/* The library decoders, for round-trip verification of generated corpora.
 * Returns 1 and the produced size on success, 0 on a library reject. */
int32_t wl_decompress(int32_t codec, const uint8_t* src, uint64_t n,
                      uint8_t* dst, uint64_t cap, uint64_t* produced) {
    pthread_once(&g_codecs_once, load_codecs);
    *produced = 0;
    if (codec == WL_CODEC_LZ4) {
        if (!g_codecs.lz4_decompress || n > 0x7FFFFFFFull || cap > 0x7FFFFFFFull) {
            return 0;
        }
        const int w = g_codecs.lz4_decompress((const char*)src, (char*)dst,
                                              (int)n, (int)cap);
        if (w < 0) {
            return 0;
        }
        *produced = (uint64_t)w;
        return 1;
    }
    if (codec == WL_CODEC_SNAPPY && g_codecs.snappy_uncompress) {
        size_t out = (size_t)cap;
        if (g_codecs.snappy_uncompress((const char*)src, (size_t)n, (char*)dst,
                                       &out) != 0) {
            return 0;
        }
        *produced = (uint64_t)out;
        return 1;
    }
    if (codec == WL_CODEC_ZSTD && g_codecs.zstd_decompress) {
        const size_t w = g_codecs.zstd_decompress(dst, (size_t)cap, src, (size_t)n);
        if (g_codecs.zstd_is_error(w)) {
            return 0;
        }
        *produced = (uint64_t)w;
        return 1;
    }
    return 0;
}

typedef struct {
    int32_t codec;
    const uint8_t* src;
    const uint64_t* src_off;
    const uint64_t* src_size;
    uint8_t* dst;
    const uint64_t* dst_off;
    const uint64_t* dst_cap;
    uint64_t* dst_size;
    _Atomic uint64_t failures;
} comp_ctx;

static void comp_one(void* p, uint64_t i) {
    comp_ctx* c = (comp_ctx*)p;
    const uint64_t w = wl_compress(c->codec, c->src + c->src_off[i],
                                   c->src_size[i], c->dst + c->dst_off[i],
                                   c->dst_cap[i]);
    c->dst_size[i] = w;
    if (w == 0 && c->src_size[i] != 0) {
        atomic_fetch_add(&c->failures, 1);
    }
}

/* Compresses n chunks in parallel; returns the number of failures. */
uint64_t wl_compress_batch(int32_t codec, const uint8_t* src,
                           const uint64_t* src_off, const uint64_t* src_size,
                           uint8_t* dst, const uint64_t* dst_off,
                           const uint64_t* dst_cap, uint64_t* dst_size,
                           uint64_t n, int32_t threads) {
    comp_ctx c = {codec, src, src_off, src_size, dst, dst_off, dst_cap, dst_size};
    atomic_init(&c.failures, 0);
    run_pool(n, threads, comp_one, &c);
    return atomic_load(&c.failures);
}

/* Moves variable-size records from bound-spaced slots into a packed arena
 * (memmove-safe when dst == src and offsets only shrink). */
void wl_pack(uint8_t* dst, const uint64_t* dst_off, const uint8_t* src,
             const uint64_t* src_off, const uint64_t* size, uint64_t n) {
    for (uint64_t i = 0; i < n; i++) {
        memmove(dst + dst_off[i], src + src_off[i], size[i]);
    }
}

/* FNV-1a 64 over a byte range, seeded; chained calls digest a corpus. */
uint64_t wl_fnv1a64(uint64_t hash, const uint8_t* data, uint64_t n) {
    for (uint64_t i = 0; i < n; i++) {
        hash ^= data[i];
        hash *= 1099511628211ull;
    }
    return hash;
}

/* ------------------------------------------------------- CPU baseline */

/* Decoders the CPU baseline times. The twins are the reference's host
 * ladders compiled from its source (or their C restatement), passed in as
 * a function pointer; the libraries are the ones loaded above. */
enum {
    WL_DEC_LIBLZ4 = 0,      /* LZ4_decompress_safe */
    WL_DEC_LIBSNAPPY = 1,   /* snappy_uncompress (capacity checked first) */
    WL_DEC_TWIN_LZ4 = 2,    /* int32 fn(src, size, dst, cap, uint64* written) */
    WL_DEC_TWIN_SNAPPY = 3, /* int32 fn(src, size, dst, cap, uint64* written, int32* rung) */
    WL_DEC_LIBZSTD = 4,     /* ZSTD_decompress (one frame per chunk) */
    WL_DEC_TWIN_ZSTD = 5,   /* int32 fn(src, size, dst, cap, uint64* written, int32* stage) */
};
typedef int32_t (*twin_lz4_fn)(const uint8_t*, uint64_t, uint8_t*, uint64_t, uint64_t*);
typedef int32_t (*twin_snappy_fn)(const uint8_t*, uint64_t, uint8_t*, uint64_t, uint64_t*,
                                  int32_t*);
typedef int (*snappy_length_fn)(const char*, size_t, size_t*);

typedef struct {
    int32_t kind;
    void* fn;
    const uint8_t* const* srcs;
    const uint64_t* sizes;
    uint8_t* const* dsts;
    const uint64_t* caps;
    uint64_t begin, end;
    int32_t* status;   /* 0 OK, else nonzero */
    uint64_t* written;
    int32_t passes;
    atomic_int* go;    /* passes released so far */
    atomic_int* done;  /* partitions finished, over all passes */
} time_part;

static snappy_length_fn g_snappy_length;

static void decode_part(time_part* t) {
    for (uint64_t i = t->begin; i < t->end; i++) {
        uint64_t w = 0;
        int32_t st = 0;
        switch (t->kind) {
            case WL_DEC_LIBLZ4: {
                const int r = g_codecs.lz4_decompress((const char*)t->srcs[i], (char*)t->dsts[i],
                                                      (int)t->sizes[i], (int)t->caps[i]);
                st = r < 0 ? 1 : 0;
                w = r < 0 ? 0 : (uint64_t)r;
                break;
            }
            case WL_DEC_LIBSNAPPY: {
                size_t len = 0;
                if (g_snappy_length((const char*)t->srcs[i], (size_t)t->sizes[i], &len) != 0 ||
                    len > t->caps[i]) {
                    st = 1;
                    break;
                }
                len = (size_t)t->caps[i];
                st = g_codecs.snappy_uncompress((const char*)t->srcs[i], (size_t)t->sizes[i],
                                                (char*)t->dsts[i], &len) != 0;
                w = st ? 0 : (uint64_t)len;
                break;
            }
            case WL_DEC_TWIN_LZ4:
                st = ((twin_lz4_fn)t->fn)(t->srcs[i], t->sizes[i], t->dsts[i], t->caps[i], &w);
                break;
            case WL_DEC_LIBZSTD: {
                const size_t r = g_codecs.zstd_decompress(t->dsts[i], (size_t)t->caps[i],
                                                          t->srcs[i], (size_t)t->sizes[i]);
                st = g_codecs.zstd_is_error(r) ? 1 : 0;
                w = st ? 0 : (uint64_t)r;
                break;
            }
            default: {
                int32_t rung = 0;
                st = ((twin_snappy_fn)t->fn)(t->srcs[i], t->sizes[i], t->dsts[i], t->caps[i], &w,
                                             &rung);
                break;
            }
        }
        if (t->status) {
            t->status[i] = st;
        }
        if (t->written) {
            t->written[i] = st == 0 ? w : 0;
        }
    }
}

/* A worker waits for each pass to be released (spinning: the wait sits
 * between passes, outside the decode), decodes its partition and counts
 * itself done. */
static void* time_worker(void* arg) {
    time_part* t = (time_part*)arg;
    for (int32_t p = 0; p < t->passes; p++) {
        while (atomic_load_explicit(t->go, memory_order_acquire) <= p) {
        }
        decode_part(t);
        atomic_fetch_add_explicit(t->done, 1, memory_order_release);
    }
    return NULL;
}

static double now_s(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* Times `passes` passes of one decoder over a batch of n chunks on `threads`
 * threads (static contiguous partition; the threads are started before the
 * first pass and wait between passes, so a pass times decode calls only,
 * into pre-sized buffers). secs[p] receives pass p's wall time; status and
 * written (either may be NULL) receive every chunk's outcome (0 = OK) of
 * the last pass. A partition whose thread could not start is decoded by
 * the calling thread. Returns the number of threads that ran, or -1 when
 * the decoder's library is absent. */
int32_t wl_time_decode(int32_t kind, void* fn, const uint8_t* const* srcs,
                       const uint64_t* sizes, uint8_t* const* dsts, const uint64_t* caps,
                       uint64_t n, int32_t threads, int32_t passes, double* secs,
                       int32_t* status, uint64_t* written) {
    pthread_once(&g_codecs_once, load_codecs);
    if (kind == WL_DEC_LIBLZ4 && !g_codecs.lz4_decompress) {
        return -1;
    }
    if (kind == WL_DEC_LIBSNAPPY) {
        if (!g_codecs.snappy_uncompress || !g_codecs.snappy) {
            return -1;
        }
        g_snappy_length =
            (snappy_length_fn)dlsym(g_codecs.snappy, "snappy_uncompressed_length");
        if (!g_snappy_length) {
            return -1;
        }
    }
    if (kind == WL_DEC_LIBZSTD && (!g_codecs.zstd_decompress || !g_codecs.zstd_is_error)) {
        return -1;
    }
    if ((kind == WL_DEC_TWIN_LZ4 || kind == WL_DEC_TWIN_SNAPPY || kind == WL_DEC_TWIN_ZSTD) &&
        !fn) {
        return -1;
    }
    if (threads < 1) {
        threads = 1;
    }
    if ((uint64_t)threads > n && n > 0) {
        threads = (int32_t)n;
    }
    time_part* parts = (time_part*)calloc((size_t)threads, sizeof(time_part));
    pthread_t* tid = (pthread_t*)calloc((size_t)threads, sizeof(pthread_t));
    int* started = (int*)calloc((size_t)threads, sizeof(int));
    if (!parts || !tid || !started) {
        free(parts);
        free(tid);
        free(started);
        return -1;
    }
    atomic_int go, done;
    atomic_init(&go, 0);
    atomic_init(&done, 0);
    int32_t running = 0;
    for (int32_t t = 0; t < threads; t++) {
        time_part* p = &parts[t];
        p->kind = kind;
        p->fn = fn;
        p->srcs = srcs;
        p->sizes = sizes;
        p->dsts = dsts;
        p->caps = caps;
        p->begin = n * (uint64_t)t / (uint64_t)threads;
        p->end = n * (uint64_t)(t + 1) / (uint64_t)threads;
        p->status = status;
        p->written = written;
        p->passes = passes;
        p->go = &go;
        p->done = &done;
        /* partition 0 is the calling thread's */
        started[t] = t != 0 && pthread_create(&tid[t], NULL, time_worker, p) == 0;
        running += started[t];
    }
    for (int32_t p = 0; p < passes; p++) {
        const double t0 = now_s();
        atomic_store_explicit(&go, p + 1, memory_order_release);
        for (int32_t t = 0; t < threads; t++) {
            if (!started[t]) {
                decode_part(&parts[t]);
            }
        }
        while (atomic_load_explicit(&done, memory_order_acquire) < (p + 1) * running) {
        }
        secs[p] = now_s() - t0;
    }
    for (int32_t t = 0; t < threads; t++) {
        if (started[t]) {
            pthread_join(tid[t], NULL);
        }
    }
    free(parts);
    free(tid);
    free(started);
    return running + 1;
}
