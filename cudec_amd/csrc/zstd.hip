/* The Zstandard batch decoder for gfx950: one wave per frame, frames taken
 * costliest first (the LZ4 decoder's cost queue: output plus what matches
 * add) from a device counter by a persistent grid.
 *
 * A frame's decode (zstd_decode.h) keeps its Huffman and FSE tables in the
 * wave's LDS (FrameTables, about 10 KiB), so a table built for one block is
 * there for the next block's Treeless or Repeat mode. Huffman-coded
 * literals are decoded four streams at a time, one lane a stream, into the
 * wave's 128 KiB literals slot in global scratch; the sequences are decoded
 * in order on every lane (the FSE states are one serial chain), 64 at a
 * time, and each batch is executed straight into the frame's output: short
 * literals and matches from before the batch one lane a sequence, the rest
 * wave-wide (zstd_decode.h, execute_batch).
 *
 * The content checksum of a frame is checked by a second kernel, four
 * lanes a frame (xxhash.h), over every frame that decoded and carries one:
 * one frame's XXH64 is a serial chain of stripes, so frames in parallel is
 * what fills the device. */
#include <mutex>

#include "common.h"
#include "cudec_zstd.h"
#include "xxhash.h"
#include "zstd_decode.h"

namespace cudec_amd {
namespace {

constexpr int kVerifyBlock = 256;
constexpr uint64_t kPoolKeep = 512ull << 20; /* scratch kept between launches */

struct Wave {
    static constexpr uint32_t kLanes = kWave;
    uint32_t lane;
    __device__ void sync() const { wave_order(); }
    __device__ uint32_t uniform(uint32_t v) const { return uniform32(v); }
    /* a reader every lane runs on the same stream. Its words stay in vector
     * registers: made wave-uniform, the sequence decode ran on the CU's
     * one scalar unit, which the CU's resident waves then shared, and it
     * was 29% slower (4.3e7 against 3.4e7 clocks a wave on 64 KiB frames) */
    struct Same {
        static __device__ uint64_t fix64(uint64_t v) { return v; }
    };
    /* a batch's sequence records: lane k holds record k in registers */
    struct Records {
        zstd::Rec mine;
        __device__ void put(uint32_t lane, uint32_t k, const zstd::Rec& v) {
            const bool me = lane == k; /* a select, not a branch */
            mine.d = me ? v.d : mine.d;
            mine.ll = me ? v.ll : mine.ll;
            mine.ml = me ? v.ml : mine.ml;
            mine.lit = me ? v.lit : mine.lit;
            mine.at = me ? v.at : mine.at;
        }
        __device__ zstd::Rec get(uint32_t k) const { /* k is the same on every lane */
            zstd::Rec r;
            r.d = ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(mine.d >> 32), (int)k) << 32) |
                  (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)mine.d, (int)k);
            r.ll = (uint32_t)__builtin_amdgcn_readlane((int)mine.ll, (int)k);
            r.ml = (uint32_t)__builtin_amdgcn_readlane((int)mine.ml, (int)k);
            r.lit = (uint32_t)__builtin_amdgcn_readlane((int)mine.lit, (int)k);
            r.at = (uint32_t)__builtin_amdgcn_readlane((int)mine.at, (int)k);
            return r;
        }
        __device__ zstd::Rec own(uint32_t, uint32_t) const { return mine; }
    };
};

/* What the decode leaves for the checksum pass, per frame. */
struct FrameCheck {
    uint64_t written;
    uint32_t checksum;
    uint32_t verify; /* 1: the frame decoded and carries a content checksum */
};

/* Scratch of one launch: the frame counter, the cost order, a FrameCheck
 * per frame, then a literals slot of kBlockCeiling bytes per wave of the
 * grid. */
constexpr size_t kCounterBytes = 256;

__global__ void zstd_reset(unsigned long long* counter) {
    if (threadIdx.x == 0) {
        *counter = 0;
    }
}

__global__ void __launch_bounds__(kWave)
    zstd_batch(const void* const* srcs, const size_t* sizes, void* const* dsts, const size_t* caps,
               size_t n, cudec_chunk_result* results, unsigned long long* counter,
               const uint32_t* order, FrameCheck* checks, uint8_t* lits_base) {
    __shared__ zstd::FrameTables tables_lds;
    ZL zstd::FrameTables& tables = *(ZL zstd::FrameTables*)&tables_lds;
    const Wave w{__lane_id()};
    ZG uint8_t* lits = (ZG uint8_t*)(lits_base + (size_t)blockIdx.x * zstd::kBlockCeiling);
    /* The loop body has no lane-conditional region: every lane issues the
     * atomic (lane 0 adds one, the others zero, so lane 0's old value is
     * the frame's index whatever order the lanes are applied in) and every
     * lane stores the same result. A lane-0-only atomic here let the
     * compiler's structurizer send lanes 1..63 back to the index read
     * without lane 0, which then read index 0 for ever. */
    for (;;) {
        const unsigned long long old = atomicAdd(counter, w.lane == 0 ? 1ull : 0ull);
        const uint64_t k = uniform64(old);
        if (k >= n) {
            break;
        }
        const uint64_t i = uniform32(order[k]);
        const ZG uint8_t* src = (const ZG uint8_t*)uniform_ptr(static_cast<const uint8_t*>(srcs[i]));
        ZG uint8_t* dst = (ZG uint8_t*)uniform_ptr(static_cast<uint8_t*>(dsts[i]));
        const uint64_t size = uniform64(sizes[i]);
        const uint64_t cap = uniform64(caps[i]);
        zstd::FrameEnd end{0, 0, false};
        const int st = (int)uniform32((uint32_t)zstd::decode_frame(w, src, size, dst, cap, tables,
                                                                     lits, end));
        write_result(results, i, st, st == CUDEC_OK ? uniform64(end.written) : 0);
        FrameCheck c;
        c.written = uniform64(end.written);
        c.checksum = uniform32(end.checksum);
        c.verify = st == CUDEC_OK && end.has_checksum ? 1u : 0u;
        checks[i] = c;
        wave_order();
    }
}

/* Frames that decoded and carry a content checksum: the low 32 bits of the
 * XXH64 (seed 0) of the output must equal it (RFC 8878 3.1.1). */
__global__ void __launch_bounds__(kVerifyBlock)
    zstd_verify(void* const* dsts, size_t n, const FrameCheck* checks,
                cudec_chunk_result* results) {
    const uint32_t lane = __lane_id();
    const uint32_t q = lane % xxh::kGroup;
    constexpr uint32_t kPerBlock = kVerifyBlock / xxh::kGroup;
    const size_t per_grid = (size_t)gridDim.x * kPerBlock;
    for (size_t i = (size_t)blockIdx.x * kPerBlock + threadIdx.x / xxh::kGroup;; i += per_grid) {
        const bool in_range = i < n;
        if (__ballot(in_range) == 0) {
            break;
        }
        const bool live = in_range && checks[i].verify != 0;
        const uint64_t h = xxh::xxh64_group(live ? (gcu8*)dsts[i] : nullptr,
                                            live ? checks[i].written : 0, 0, q, lane - q);
        if (live && q == 0 && (uint32_t)h != checks[i].checksum) {
            write_result(results, i, CUDEC_ERR_CORRUPT_INPUT, 0);
        }
    }
}

std::mutex g_pool_lock;
constexpr int kMaxDevices = 64;
hipMemPool_t g_pool[kMaxDevices];
bool g_pool_tried[kMaxDevices];

/* The device's scratch pool for this decoder, made on first use; it keeps
 * up to kPoolKeep bytes of freed scratch between launches. */
hipMemPool_t zstd_pool(int dev) {
    if (dev < 0 || dev >= kMaxDevices) {
        return nullptr;
    }
    std::lock_guard<std::mutex> hold(g_pool_lock);
    if (!g_pool_tried[dev]) {
        g_pool_tried[dev] = true;
        hipMemPoolProps props = {};
        props.allocType = hipMemAllocationTypePinned;
        props.location.type = hipMemLocationTypeDevice;
        props.location.id = dev;
        hipMemPool_t pool = nullptr;
        if (hipMemPoolCreate(&pool, &props) == hipSuccess) {
            uint64_t keep = kPoolKeep;
            if (hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep) == hipSuccess) {
                g_pool[dev] = pool;
            } else {
                (void)hipMemPoolDestroy(pool);
            }
        }
        (void)hipGetLastError(); /* a pool that could not be made is not an error */
    }
    return g_pool[dev];
}

}  // namespace

/* tuned.hip: the batch's cost order. */
size_t cost_order_bytes(size_t n);
hipError_t cost_order(const size_t* sizes, const size_t* caps, size_t n, void* work,
                      hipStream_t stream, const uint32_t** order);

hipError_t launch_zstd(const void* const* srcs, const size_t* sizes, void* const* dsts,
                       const size_t* caps, size_t n, cudec_chunk_result* results,
                       hipStream_t stream) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) {
        return e;
    }
    int cus = 0;
    e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e != hipSuccess) {
        return e;
    }
    /* as many one-wave blocks as are resident at once (LDS and registers
     * decide), on every CU */
    int per_cu = 0;
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, zstd_batch, kWave, 0);
    if (e != hipSuccess) {
        return e;
    }
    const size_t waves = (size_t)(cus > 0 ? cus : 1) * (size_t)(per_cu > 0 ? per_cu : 1);
    const size_t grid = n < waves ? n : waves;
    if (n > UINT32_MAX) {
        return hipErrorInvalidValue; /* the cost order indexes frames in 32 bits */
    }
    const size_t order_bytes = (cost_order_bytes(n) + 255) & ~(size_t)255;
    const size_t checks_bytes = (sizeof(FrameCheck) * n + 255) & ~(size_t)255;
    const size_t bytes = kCounterBytes + order_bytes + checks_bytes + grid * zstd::kBlockCeiling;
    hipStreamCaptureStatus capture = hipStreamCaptureStatusNone;
    e = hipStreamIsCapturing(stream, &capture);
    if (e != hipSuccess) {
        return e;
    }
    hipMemPool_t pool = capture == hipStreamCaptureStatusNone ? zstd_pool(dev) : nullptr;
    void* mem = nullptr;
    e = pool != nullptr ? hipMallocFromPoolAsync(&mem, bytes, pool, stream)
                        : hipMallocAsync(&mem, bytes, stream);
    if (e != hipSuccess) {
        return e;
    }
    uint8_t* base = static_cast<uint8_t*>(mem);
    auto* counter = reinterpret_cast<unsigned long long*>(base);
    void* order_work = base + kCounterBytes;
    auto* checks = reinterpret_cast<FrameCheck*>(base + kCounterBytes + order_bytes);
    uint8_t* lits = base + kCounterBytes + order_bytes + checks_bytes;
    const uint32_t* order = nullptr;
    hipLaunchKernelGGL(zstd_reset, dim3(1), dim3(kWave), 0, stream, counter);
    e = hipGetLastError();
    if (e == hipSuccess) {
        e = cost_order(sizes, caps, n, order_work, stream, &order);
    }
    if (e == hipSuccess) {
        hipLaunchKernelGGL(zstd_batch, dim3((unsigned)grid), dim3(kWave), 0, stream, srcs, sizes,
                           dsts, caps, n, results, counter, order, checks, lits);
        e = hipGetLastError();
    }
    if (e == hipSuccess) {
        constexpr size_t kPerBlock = kVerifyBlock / xxh::kGroup;
        size_t vgrid = (n + kPerBlock - 1) / kPerBlock;
        vgrid = vgrid > 8192 ? 8192 : vgrid;
        hipLaunchKernelGGL(zstd_verify, dim3((unsigned)vgrid), dim3(kVerifyBlock), 0, stream, dsts,
                           n, checks, results);
        e = hipGetLastError();
    }
    const hipError_t f = hipFreeAsync(mem, stream);
    return e != hipSuccess ? e : f;
}

}  // namespace cudec_amd
