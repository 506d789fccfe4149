/* The reusable streaming-decode context: LZ4 blocks from host memory,
 * decoded on the device through cudec_lz4_decompress_batch.
 *
 * Contract: include/cudec.h:140-218 (iderex/cudec's src/stream.cpp:
 * ValidateStreamArgs 196-216, StampNotDecoded 218-230, DecodeStreamCtx
 * 238-488, the entry points 492-593, CopyWaveToHost 163-184).
 *
 * This implementation's shape: the call is split into waves sized by a
 * staging budget charged to its largest chunk (so a hostile size mix cannot
 * inflate staging). Staging is double-buffered -- two pinned source slots,
 * two device source slots, and for host output two device destination and
 * two pinned destination slots -- so the host-side pack of wave w+1 and the
 * host-side unpack of wave w-1 overlap the device decode of wave w on the
 * context's one stream. Host output comes back as one asynchronous copy per
 * wave into pinned memory, then exactly bytes_written bytes per chunk go to
 * the caller. Everything grows on demand and is kept for the next call. */
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <new>

#include "cudec.h"
#include "hip_owner.h"

namespace cudec_amd {

/* One chunk per wave and at most INT32_MAX waves (src/batch_limits.h:28 for
 * wave64); the batch entry applies the same bound. */
constexpr size_t kStreamMaxChunks = static_cast<size_t>(INT32_MAX) * 64;
/* Staging charged per wave (pinned + device source, device + pinned
 * destination on the host-output path), against the largest chunk. */
constexpr size_t kWaveBudget = size_t{1} << 30;
constexpr size_t kWaveChunksMax = 16384;

struct Slot {
    PinnedRegion p_src;
    DeviceRegion d_src;
    PinnedRegion p_meta;
    DeviceRegion d_meta;
    DeviceRegion d_dst;
    PinnedRegion p_dst;
    Event staged;  /* the H2D of this slot's pinned source has completed */
    Event landed;  /* this slot's results and output are in pinned memory */
    bool staged_pending = false;
    bool landed_pending = false;
    size_t begin = 0, count = 0;
};

}  // namespace cudec_amd

struct cudec_stream_ctx {
    cudec_amd::Stream stream;
    cudec_amd::Slot slot[2];
    cudec_amd::DeviceRegion d_results;
    cudec_amd::PinnedRegion p_results;
    bool poisoned = false;
};

namespace cudec_amd {
namespace {

cudec_status validate_stream_args(const void* const* h_src_ptrs, const size_t* h_src_sizes,
                                  void* const* dst_ptrs, const size_t* dst_caps,
                                  size_t chunk_count, cudec_mem_space dst_space,
                                  const cudec_chunk_result* h_results) {
    if (h_src_ptrs == nullptr || h_src_sizes == nullptr || dst_ptrs == nullptr ||
        dst_caps == nullptr || h_results == nullptr || chunk_count == 0 ||
        chunk_count > kStreamMaxChunks ||
        (dst_space != CUDEC_MEM_HOST && dst_space != CUDEC_MEM_DEVICE)) {
        return CUDEC_ERR_INVALID_ARGUMENT;
    }
    for (size_t i = 0; i < chunk_count; i++) {
        if ((h_src_ptrs[i] == nullptr && h_src_sizes[i] != 0) ||
            (dst_ptrs[i] == nullptr && dst_caps[i] != 0)) {
            return CUDEC_ERR_INVALID_ARGUMENT;
        }
    }
    return CUDEC_OK;
}

void stamp(cudec_chunk_result* r, size_t n, cudec_status st) {
    for (size_t i = 0; i < n; i++) {
        r[i].status = static_cast<int32_t>(st);
        r[i].reserved = 0;
        r[i].bytes_written = 0;
    }
}

size_t choose_wave(const size_t* src_sizes, const size_t* caps, size_t n, bool host_out) {
    size_t biggest = 1;
    for (size_t i = 0; i < n; i++) {
        size_t b = src_sizes[i];
        if (host_out) {
            b = SIZE_MAX - b < caps[i] ? SIZE_MAX : b + caps[i];
        }
        if (b > biggest) {
            biggest = b;
        }
    }
    size_t wave = kWaveBudget / biggest;
    if (wave == 0) {
        wave = 1;
    }
    if (wave > kWaveChunksMax) {
        wave = kWaveChunksMax;
    }
    return wave > n ? n : wave;
}

/* Unpacks one landed host-output wave: exactly bytes_written bytes per chunk
 * into the caller's buffer. A device-reported length beyond the chunk's
 * capacity is a device fault, never a copy length (src/stream.cpp:163-184). */
cudec_status unpack_wave(const Slot& s, cudec_chunk_result* pres, void* const* dst_ptrs,
                         const size_t* caps) {
    const unsigned char* base = static_cast<const unsigned char*>(s.p_dst.data());
    size_t off = 0;
    for (size_t j = 0; j < s.count; j++) {
        const size_t i = s.begin + j;
        cudec_chunk_result& r = pres[i];
        const size_t bw = static_cast<size_t>(r.bytes_written);
        if (bw > caps[i]) {
            r.status = static_cast<int32_t>(CUDEC_ERR_CUDA);
            r.reserved = 0;
            r.bytes_written = 0;
            return CUDEC_ERR_CUDA;
        }
        if (r.status == CUDEC_OK && bw != 0) {
            std::memcpy(dst_ptrs[i], base + off, bw);
        }
        off += caps[i];
    }
    return CUDEC_OK;
}

cudec_status decode_stream(cudec_stream_ctx& ctx, const void* const* h_src_ptrs,
                           const size_t* h_src_sizes, void* const* dst_ptrs,
                           const size_t* dst_caps, size_t n, cudec_mem_space space,
                           cudec_chunk_result* h_results) {
    const bool host_out = space == CUDEC_MEM_HOST;
    stamp(h_results, n, CUDEC_ERR_CUDA);
    const size_t wave = choose_wave(h_src_sizes, dst_caps, n, host_out);
    const size_t waves = (n + wave - 1) / wave;
    size_t max_src = 1, max_dst = 1;
    for (size_t w = 0; w < waves; w++) {
        const size_t b = w * wave, e = b + wave < n ? b + wave : n;
        size_t ws = 0, wd = 0;
        for (size_t i = b; i < e; i++) {
            if (SIZE_MAX - ws < h_src_sizes[i]) {
                return CUDEC_ERR_CORRUPT_INPUT;
            }
            ws += h_src_sizes[i];
            if (host_out) {
                if (SIZE_MAX - wd < dst_caps[i]) {
                    return CUDEC_ERR_CORRUPT_INPUT;
                }
                wd += dst_caps[i];
            }
        }
        max_src = ws > max_src ? ws : max_src;
        max_dst = wd > max_dst ? wd : max_dst;
    }
    if (n > SIZE_MAX / sizeof(cudec_chunk_result)) {
        return CUDEC_ERR_CORRUPT_INPUT;
    }
    const size_t meta_stride = wave * sizeof(void*);
    const size_t meta_bytes = 4 * meta_stride;
    const size_t res_bytes = n * sizeof(cudec_chunk_result);
    bool ok = true;
    for (int k = 0; k < 2 && ok; k++) {
        Slot& s = ctx.slot[k];
        ok = s.p_src.reserve(max_src) == hipSuccess && s.d_src.reserve(max_src) == hipSuccess &&
             s.p_meta.reserve(meta_bytes) == hipSuccess &&
             s.d_meta.reserve(meta_bytes) == hipSuccess &&
             (!host_out || (s.d_dst.reserve(max_dst) == hipSuccess &&
                            s.p_dst.reserve(max_dst) == hipSuccess));
        s.staged_pending = false;
        s.landed_pending = false;
    }
    ok = ok && ctx.d_results.reserve(res_bytes) == hipSuccess &&
         ctx.p_results.reserve(res_bytes) == hipSuccess;
    if (!ok) {
        ctx.poisoned = true;
        return CUDEC_ERR_CUDA;
    }
    cudec_chunk_result* pres = static_cast<cudec_chunk_result*>(ctx.p_results.data());
    stamp(pres, n, CUDEC_ERR_CUDA);
    hipStream_t st = ctx.stream.get();
    cudec_status failure = CUDEC_OK;

    for (size_t w = 0; w < waves && failure == CUDEC_OK; w++) {
        Slot& s = ctx.slot[w & 1];
        /* the slot's previous wave: its pinned source must be uploaded, and
         * on the host-output path its output unpacked, before reuse */
        if (s.staged_pending) {
            if (hipEventSynchronize(s.staged.get()) != hipSuccess) {
                failure = CUDEC_ERR_CUDA;
                break;
            }
            s.staged_pending = false;
        }
        if (s.landed_pending) {
            if (hipEventSynchronize(s.landed.get()) != hipSuccess) {
                failure = CUDEC_ERR_CUDA;
                break;
            }
            s.landed_pending = false;
            if (host_out) {
                failure = unpack_wave(s, pres, dst_ptrs, dst_caps);
                if (failure != CUDEC_OK) {
                    break;
                }
            }
        }
        s.begin = w * wave;
        s.count = (s.begin + wave < n ? s.begin + wave : n) - s.begin;
        unsigned char* ps = static_cast<unsigned char*>(s.p_src.data());
        unsigned char* ds = static_cast<unsigned char*>(s.d_src.data());
        unsigned char* dd = static_cast<unsigned char*>(s.d_dst.data());
        unsigned char* pm = static_cast<unsigned char*>(s.p_meta.data());
        const void** m_src = reinterpret_cast<const void**>(pm);
        size_t* m_size = reinterpret_cast<size_t*>(pm + meta_stride);
        void** m_dst = reinterpret_cast<void**>(pm + 2 * meta_stride);
        size_t* m_cap = reinterpret_cast<size_t*>(pm + 3 * meta_stride);
        size_t so = 0, dof = 0;
        for (size_t j = 0; j < s.count; j++) {
            const size_t i = s.begin + j;
            if (h_src_sizes[i] != 0) {
                std::memcpy(ps + so, h_src_ptrs[i], h_src_sizes[i]);
            }
            m_src[j] = ds + so;
            m_size[j] = h_src_sizes[i];
            m_cap[j] = dst_caps[i];
            if (host_out) {
                m_dst[j] = dd + dof;
                dof += dst_caps[i];
            } else {
                m_dst[j] = dst_ptrs[i];
            }
            so += h_src_sizes[i];
        }
        unsigned char* dm = static_cast<unsigned char*>(s.d_meta.data());
        cudec_chunk_result* dres = static_cast<cudec_chunk_result*>(ctx.d_results.data()) + s.begin;
        if ((so != 0 && hipMemcpyAsync(ds, ps, so, hipMemcpyHostToDevice, st) != hipSuccess) ||
            hipMemcpyAsync(dm, pm, meta_bytes, hipMemcpyHostToDevice, st) != hipSuccess ||
            hipEventRecord(s.staged.get(), st) != hipSuccess) {
            failure = CUDEC_ERR_CUDA;
            break;
        }
        s.staged_pending = true;
        const cudec_status launched = cudec_lz4_decompress_batch(
            reinterpret_cast<const void* const*>(dm),
            reinterpret_cast<const size_t*>(dm + meta_stride),
            reinterpret_cast<void* const*>(dm + 2 * meta_stride),
            reinterpret_cast<const size_t*>(dm + 3 * meta_stride), s.count, dres,
            reinterpret_cast<cudec_stream_t>(st));
        if (launched != CUDEC_OK) {
            failure = launched;
            break;
        }
        if (hipMemcpyAsync(pres + s.begin, dres, s.count * sizeof(cudec_chunk_result),
                           hipMemcpyDeviceToHost, st) != hipSuccess ||
            (host_out && dof != 0 &&
             hipMemcpyAsync(s.p_dst.data(), dd, dof, hipMemcpyDeviceToHost, st) != hipSuccess) ||
            hipEventRecord(s.landed.get(), st) != hipSuccess) {
            failure = CUDEC_ERR_CUDA;
            break;
        }
        s.landed_pending = true;
    }
    /* drain: nothing of this call outlives it */
    if (hipStreamSynchronize(st) != hipSuccess && failure == CUDEC_OK) {
        failure = CUDEC_ERR_CUDA;
    }
    if (hipGetLastError() != hipSuccess && failure == CUDEC_OK) {
        failure = CUDEC_ERR_CUDA;
    }
    /* unpack the waves still in flight, in order */
    for (size_t w = (waves >= 2 ? waves - 2 : 0); w < waves && failure == CUDEC_OK; w++) {
        Slot& s = ctx.slot[w & 1];
        if (s.landed_pending && s.begin == w * wave) {
            s.landed_pending = false;
            if (host_out) {
                failure = unpack_wave(s, pres, dst_ptrs, dst_caps);
            }
        }
    }
    for (int k = 0; k < 2; k++) {
        ctx.slot[k].staged_pending = false;
        ctx.slot[k].landed_pending = false;
    }
    std::memcpy(h_results, pres, res_bytes);
    if (failure == CUDEC_ERR_CUDA) {
        ctx.poisoned = true;
        return failure;
    }
    if (failure != CUDEC_OK) {
        return failure;
    }
    for (size_t i = 0; i < n; i++) {
        if (h_results[i].status != CUDEC_OK) {
            return static_cast<cudec_status>(h_results[i].status);
        }
    }
    return CUDEC_OK;
}

}  // namespace
}  // namespace cudec_amd

using namespace cudec_amd;

extern "C" {

cudec_status cudec_stream_ctx_create(cudec_stream_ctx** out_ctx) {
    if (out_ctx == nullptr) {
        return CUDEC_ERR_INVALID_ARGUMENT;
    }
    *out_ctx = nullptr;
    cudec_stream_ctx* ctx = new (std::nothrow) cudec_stream_ctx();
    if (ctx == nullptr) {
        return CUDEC_ERR_CUDA;
    }
    if (make_stream(ctx->stream) != hipSuccess ||
        make_sync_event(ctx->slot[0].staged) != hipSuccess ||
        make_sync_event(ctx->slot[0].landed) != hipSuccess ||
        make_sync_event(ctx->slot[1].staged) != hipSuccess ||
        make_sync_event(ctx->slot[1].landed) != hipSuccess) {
        delete ctx;
        return CUDEC_ERR_CUDA;
    }
    *out_ctx = ctx;
    return CUDEC_OK;
}

cudec_status cudec_lz4_decompress_stream_ctx(cudec_stream_ctx* ctx,
                                             const void* const* h_src_ptrs,
                                             const size_t* h_src_sizes, void* const* dst_ptrs,
                                             const size_t* dst_caps, size_t chunk_count,
                                             cudec_mem_space dst_space,
                                             cudec_chunk_result* h_results) {
    if (ctx == nullptr) {
        return CUDEC_ERR_INVALID_ARGUMENT;
    }
    const cudec_status args = validate_stream_args(h_src_ptrs, h_src_sizes, dst_ptrs, dst_caps,
                                                   chunk_count, dst_space, h_results);
    if (args != CUDEC_OK) {
        return args;
    }
    if (ctx->poisoned) {
        stamp(h_results, chunk_count, CUDEC_ERR_CUDA);
        return CUDEC_ERR_CUDA;
    }
    try {
        return decode_stream(*ctx, h_src_ptrs, h_src_sizes, dst_ptrs, dst_caps, chunk_count,
                             dst_space, h_results);
    } catch (...) {
        ctx->poisoned = true;
        return CUDEC_ERR_CUDA;
    }
}

void cudec_stream_ctx_destroy(cudec_stream_ctx* ctx) {
    if (ctx == nullptr) {
        return;
    }
    if (ctx->stream.get() != nullptr) {
        (void)hipStreamSynchronize(ctx->stream.get());
    }
    delete ctx;
}

cudec_status cudec_lz4_decompress_stream(const void* const* h_src_ptrs,
                                         const size_t* h_src_sizes, void* const* dst_ptrs,
                                         const size_t* dst_caps, size_t chunk_count,
                                         cudec_mem_space dst_space,
                                         cudec_chunk_result* h_results) {
    const cudec_status args = validate_stream_args(h_src_ptrs, h_src_sizes, dst_ptrs, dst_caps,
                                                   chunk_count, dst_space, h_results);
    if (args != CUDEC_OK) {
        return args;
    }
    cudec_stream_ctx* ctx = nullptr;
    const cudec_status created = cudec_stream_ctx_create(&ctx);
    if (created != CUDEC_OK) {
        return created;
    }
    const cudec_status st = cudec_lz4_decompress_stream_ctx(
        ctx, h_src_ptrs, h_src_sizes, dst_ptrs, dst_caps, chunk_count, dst_space, h_results);
    cudec_stream_ctx_destroy(ctx);
    return st;
}

}  // extern "C"
