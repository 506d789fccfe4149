/* cudec_lz4f_decompress: one LZ4 frame (the .lz4 container), host memory
 * to host memory, through the batch block decoder.
 *
 * Contract: include/cudec.h:110-132; behaviour follows iderex/cudec's
 * src/frame.cpp:282-422 (parse, status mapping) and 53-278 (decode and
 * assembly). Format: the public LZ4 frame format (magic 0x184D2204, FLG,
 * BD, optional content size, header checksum byte, data blocks with the
 * high bit marking an uncompressed block, optional block checksums, end
 * mark, optional content checksum; checksums are XXH32 with seed 0).
 *
 * Decoding here: every compressed block of the frame goes to the device in
 * one upload and one batch launch (one destination slot of the frame's
 * block maximum per block), and the decoded blocks come back with one copy
 * per run of blocks that are contiguous both on the device and in the
 * output -- every block but the last decodes to exactly the block maximum,
 * so a frame of compressed blocks comes back in one copy. The caller's
 * capacity is checked before any output byte is written.
 *
 * Checksums: block checksums (XXH32 of each block's stored bytes) are
 * computed on the device, every block in parallel, from the same upload
 * (stored blocks are uploaded too when the frame carries them), and
 * checked in block order before any decode result is used. The header
 * checksum (a few bytes) and the content checksum are computed on the
 * host: the content checksum is one XXH32 stream over the whole output,
 * and its stripes form a serial chain, which one host core hashes at
 * about 13 GB/s and one group of GPU lanes at about 0.5 GB/s. */
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <vector>

#include "cudec.h"
#include "cudec_mi355x.h"
#include "hip_owner.h"

namespace cudec_amd {
namespace {

constexpr uint32_t kFrameMagic = 0x184D2204u;

uint32_t le32(const unsigned char* p) {
    uint32_t v;
    std::memcpy(&v, p, 4);
    return v;
}

uint64_t le64(const unsigned char* p) {
    uint64_t v;
    std::memcpy(&v, p, 8);
    return v;
}

/* XXH32, seed 0, from the published xxHash specification: four 32-bit
 * accumulators over 16-byte stripes, then 4-byte and 1-byte tail rounds
 * and the final avalanche. */
constexpr uint32_t kP1 = 0x9E3779B1u, kP2 = 0x85EBCA77u, kP3 = 0xC2B2AE3Du,
                   kP4 = 0x27D4EB2Fu, kP5 = 0x165667B1u;

inline uint32_t rotl32(uint32_t x, int r) { return (x << r) | (x >> (32 - r)); }

uint32_t xxh32(const unsigned char* p, size_t len) {
    const unsigned char* const end = p + len;
    uint32_t h;
    if (len >= 16) {
        uint32_t a = kP1 + kP2, b = kP2, c = 0, d = 0u - kP1;
        const unsigned char* const last = end - 16;
        do {
            a = rotl32(a + le32(p) * kP2, 13) * kP1;
            b = rotl32(b + le32(p + 4) * kP2, 13) * kP1;
            c = rotl32(c + le32(p + 8) * kP2, 13) * kP1;
            d = rotl32(d + le32(p + 12) * kP2, 13) * kP1;
            p += 16;
        } while (p <= last);
        h = rotl32(a, 1) + rotl32(b, 7) + rotl32(c, 12) + rotl32(d, 18);
    } else {
        h = kP5;
    }
    h += static_cast<uint32_t>(len);
    while (p + 4 <= end) {
        h = rotl32(h + le32(p) * kP3, 17) * kP4;
        p += 4;
    }
    while (p < end) {
        h = rotl32(h + static_cast<uint32_t>(*p) * kP5, 11) * kP1;
        p++;
    }
    h ^= h >> 15;
    h *= kP2;
    h ^= h >> 13;
    h *= kP3;
    h ^= h >> 16;
    return h;
}

struct Block {
    size_t off;
    size_t len;
    bool stored;  /* the high bit: the block is stored uncompressed */
    uint32_t sum; /* the block checksum the frame carries (when it does) */
};

/* Checks the block checksums on the device (when `sums`), decodes the
 * compressed blocks in one batch, then assembles the output. */
cudec_status decode_blocks(const unsigned char* f, const std::vector<Block>& blocks,
                           size_t block_max, bool sums, unsigned char* out, size_t cap,
                           size_t* total_out) {
    /* the blocks that go to the device: the compressed ones, and every one
     * when their checksums are checked there */
    std::vector<size_t> up, ci;
    size_t total_src = 0;
    for (size_t i = 0; i < blocks.size(); i++) {
        if (sums || !blocks[i].stored) {
            up.push_back(i);
            total_src += blocks[i].len;
        }
        if (!blocks[i].stored) {
            ci.push_back(i);
        }
    }
    const size_t n = ci.size(), nu = up.size();
    std::vector<size_t> decoded(blocks.size(), 0);
    Stream stream;
    DeviceRegion d_src, d_dst, d_meta, d_res, d_sums;
    if (nu != 0) {
        if (n > SIZE_MAX / block_max || nu > SIZE_MAX / (6 * sizeof(void*))) {
            return CUDEC_ERR_CORRUPT_INPUT;
        }
        /* tables: decode sources, sizes, destinations, capacities (n
         * each), then checksum sources and sizes (nu each) */
        const size_t stride = n * sizeof(void*), ustride = nu * sizeof(void*);
        if (make_stream(stream) != hipSuccess || d_src.reserve(total_src) != hipSuccess ||
            d_dst.reserve(n * block_max) != hipSuccess ||
            d_meta.reserve(4 * stride + 2 * ustride) != hipSuccess ||
            d_res.reserve(n * sizeof(cudec_chunk_result)) != hipSuccess ||
            d_sums.reserve(nu * sizeof(uint32_t)) != hipSuccess) {
            return CUDEC_ERR_CUDA;
        }
        std::vector<unsigned char> stage(total_src);
        std::vector<unsigned char> meta(4 * stride + 2 * ustride);
        unsigned char* ds = static_cast<unsigned char*>(d_src.data());
        unsigned char* dd = static_cast<unsigned char*>(d_dst.data());
        size_t so = 0, k = 0;
        for (size_t u = 0; u < nu; u++) {
            const Block& b = blocks[up[u]];
            std::memcpy(stage.data() + so, f + b.off, b.len);
            const void* sp = ds + so;
            std::memcpy(meta.data() + 4 * stride + u * sizeof(void*), &sp, sizeof(void*));
            std::memcpy(meta.data() + 4 * stride + ustride + u * sizeof(size_t), &b.len,
                        sizeof(size_t));
            if (!b.stored) {
                void* dp = dd + k * block_max;
                std::memcpy(meta.data() + k * sizeof(void*), &sp, sizeof(void*));
                std::memcpy(meta.data() + stride + k * sizeof(size_t), &b.len, sizeof(size_t));
                std::memcpy(meta.data() + 2 * stride + k * sizeof(void*), &dp, sizeof(void*));
                std::memcpy(meta.data() + 3 * stride + k * sizeof(size_t), &block_max,
                            sizeof(size_t));
                k++;
            }
            so += b.len;
        }
        hipStream_t st = stream.get();
        if (hipMemcpyAsync(ds, stage.data(), total_src, hipMemcpyHostToDevice, st) != hipSuccess ||
            hipMemcpyAsync(d_meta.data(), meta.data(), meta.size(), hipMemcpyHostToDevice, st) !=
                hipSuccess) {
            (void)hipStreamSynchronize(st);
            return CUDEC_ERR_CUDA;
        }
        unsigned char* dm = static_cast<unsigned char*>(d_meta.data());
        const cudec_stream_t cs = reinterpret_cast<cudec_stream_t>(st);
        std::vector<uint32_t> got(nu);
        cudec_status launched = CUDEC_OK;
        bool copied = true;
        if (sums) {
            launched = cudec_mi355x_xxh32_batch(
                reinterpret_cast<const void* const*>(dm + 4 * stride),
                reinterpret_cast<const size_t*>(dm + 4 * stride + ustride), nu, 0,
                static_cast<uint32_t*>(d_sums.data()), cs);
            copied = launched == CUDEC_OK &&
                     hipMemcpyAsync(got.data(), d_sums.data(), nu * sizeof(uint32_t),
                                    hipMemcpyDeviceToHost, st) == hipSuccess;
        }
        std::vector<cudec_chunk_result> res(n);
        if (launched == CUDEC_OK && n != 0) {
            launched = cudec_lz4_decompress_batch(
                reinterpret_cast<const void* const*>(dm),
                reinterpret_cast<const size_t*>(dm + stride),
                reinterpret_cast<void* const*>(dm + 2 * stride),
                reinterpret_cast<const size_t*>(dm + 3 * stride), n,
                static_cast<cudec_chunk_result*>(d_res.data()), cs);
            copied = copied && launched == CUDEC_OK &&
                     hipMemcpyAsync(res.data(), d_res.data(), n * sizeof(cudec_chunk_result),
                                    hipMemcpyDeviceToHost, st) == hipSuccess;
        }
        /* the staging vectors must outlive the asynchronous uploads */
        if (hipStreamSynchronize(st) != hipSuccess) {
            return CUDEC_ERR_CUDA;
        }
        if (launched != CUDEC_OK) {
            return launched;
        }
        if (!copied) {
            return CUDEC_ERR_CUDA;
        }
        if (sums) {
            for (size_t u = 0; u < nu; u++) {
                if (got[u] != blocks[up[u]].sum) {
                    return CUDEC_ERR_CORRUPT_INPUT;
                }
            }
        }
        for (size_t k = 0; k < n; k++) {
            /* any block the decoder refuses (including one longer than the
             * block maximum) makes the frame corrupt */
            if (res[k].status != CUDEC_OK || res[k].bytes_written > block_max) {
                return CUDEC_ERR_CORRUPT_INPUT;
            }
            decoded[ci[k]] = static_cast<size_t>(res[k].bytes_written);
        }
    }
    size_t total = 0;
    for (size_t i = 0; i < blocks.size(); i++) {
        const size_t len = blocks[i].stored ? blocks[i].len : decoded[i];
        if (len > cap - total) {
            return CUDEC_ERR_OUTPUT_TOO_SMALL;
        }
        total += len;
    }
    size_t run_dev = 0, run_out = 0, run_len = 0, pos = 0, k = 0;
    hipError_t err = hipSuccess;
    for (size_t i = 0; i < blocks.size() && err == hipSuccess; i++) {
        if (blocks[i].stored) {
            if (blocks[i].len != 0) {
                std::memcpy(out + pos, f + blocks[i].off, blocks[i].len);
            }
            pos += blocks[i].len;
            continue;
        }
        const size_t len = decoded[i];
        const size_t dev = k * block_max;
        if (run_len != 0 && run_dev + run_len == dev && run_out + run_len == pos) {
            run_len += len;
        } else {
            if (run_len != 0) {
                err = hipMemcpyAsync(out + run_out, static_cast<unsigned char*>(d_dst.data()) + run_dev,
                                     run_len, hipMemcpyDeviceToHost, stream.get());
            }
            run_dev = dev;
            run_out = pos;
            run_len = len;
        }
        pos += len;
        k++;
    }
    if (err == hipSuccess && run_len != 0) {
        err = hipMemcpyAsync(out + run_out, static_cast<unsigned char*>(d_dst.data()) + run_dev, run_len,
                             hipMemcpyDeviceToHost, stream.get());
    }
    if (n != 0 && hipStreamSynchronize(stream.get()) != hipSuccess) {
        err = hipErrorUnknown;
    }
    if (err != hipSuccess) {
        return CUDEC_ERR_CUDA;
    }
    *total_out = total;
    return CUDEC_OK;
}

cudec_status decode_frame(const unsigned char* f, size_t size, unsigned char* out, size_t cap,
                          size_t* written) {
    if (size < 7 || le32(f) != kFrameMagic) {
        return CUDEC_ERR_CORRUPT_INPUT;
    }
    const unsigned flg = f[4], bd = f[5];
    if (((flg >> 6) & 3) != 1 || (flg & 2) != 0 || (bd & 0x8F) != 0) {
        return CUDEC_ERR_CORRUPT_INPUT; /* version, reserved bits */
    }
    const unsigned bmax = (bd >> 4) & 7;
    if (bmax < 4) {
        return CUDEC_ERR_CORRUPT_INPUT;
    }
    const size_t block_max = size_t{64 << 10} << ((bmax - 4) * 2);
    const bool independent = (flg >> 5) & 1, block_sum = (flg >> 4) & 1,
               has_size = (flg >> 3) & 1, content_sum = (flg >> 2) & 1, dict = flg & 1;
    if (!independent || dict) {
        return CUDEC_ERR_UNSUPPORTED;
    }
    size_t pos = 6;
    uint64_t declared = 0;
    if (has_size) {
        if (size < pos + 8) {
            return CUDEC_ERR_CORRUPT_INPUT;
        }
        declared = le64(f + pos);
        pos += 8;
    }
    if (size < pos + 1) {
        return CUDEC_ERR_CORRUPT_INPUT;
    }
    if (((xxh32(f + 4, pos - 4) >> 8) & 0xFF) != f[pos]) {
        return CUDEC_ERR_CORRUPT_INPUT;
    }
    pos++;
    std::vector<Block> blocks;
    uint64_t fuel = size / 4 + 1; /* each block consumes at least 4 bytes */
    bool ended = false;
    while (fuel-- != 0) {
        if (size < pos + 4) {
            return CUDEC_ERR_CORRUPT_INPUT;
        }
        const uint32_t bs = le32(f + pos);
        pos += 4;
        if (bs == 0) {
            ended = true;
            break;
        }
        const size_t len = bs & 0x7FFFFFFFu;
        if (len == 0 || len > block_max || size - pos < len) {
            return CUDEC_ERR_CORRUPT_INPUT;
        }
        uint32_t sum = 0;
        if (block_sum) {
            /* the checksum itself is checked on the device, with the others */
            if (size - pos - len < 4) {
                return CUDEC_ERR_CORRUPT_INPUT;
            }
            sum = le32(f + pos + len);
        }
        blocks.push_back(Block{pos, len, ((bs >> 31) & 1) != 0, sum});
        pos += len + (block_sum ? 4 : 0);
    }
    if (!ended) {
        return CUDEC_ERR_CORRUPT_INPUT;
    }
    size_t total = 0;
    const cudec_status st = decode_blocks(f, blocks, block_max, block_sum, out, cap, &total);
    if (st != CUDEC_OK) {
        return st;
    }
    if (has_size && declared != total) {
        return CUDEC_ERR_CORRUPT_INPUT;
    }
    if (content_sum) {
        if (size < pos + 4 || xxh32(out, total) != le32(f + pos)) {
            return CUDEC_ERR_CORRUPT_INPUT;
        }
    }
    *written = total;
    return CUDEC_OK;
}

}  // namespace
}  // namespace cudec_amd

extern "C" cudec_status cudec_lz4f_decompress(const void* frame, size_t frame_size, void* dst,
                                              size_t dst_capacity, size_t* bytes_written) {
    if (bytes_written != nullptr) {
        *bytes_written = 0;
    }
    if (frame == nullptr || bytes_written == nullptr || (dst == nullptr && dst_capacity != 0)) {
        return CUDEC_ERR_INVALID_ARGUMENT;
    }
    try {
        size_t w = 0;
        const cudec_status st = cudec_amd::decode_frame(static_cast<const unsigned char*>(frame),
                                                        frame_size,
                                                        static_cast<unsigned char*>(dst),
                                                        dst_capacity, &w);
        if (st == CUDEC_OK) {
            *bytes_written = w;
        }
        return st;
    } catch (...) {
        return CUDEC_ERR_CUDA;
    }
}
