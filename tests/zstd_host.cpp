/* The device Zstd decoder's frame logic (cudec_amd/csrc/zstd_decode.h)
 * built for the CPU with a one-lane wave: TEST INFRASTRUCTURE ONLY. It lets
 * the CPU suite hold that logic to libzstd and to the reference's frame
 * twin over corpora and mutants (and run it under the sanitizers) before
 * the same code runs in the kernel. The product never loads it. */
#include "zstd_decode.h"

#include <cstdlib>
#include <cstring>

namespace {

struct OneLane {
    static constexpr uint32_t kLanes = 1;
    uint32_t lane = 0;
    void sync() const {}
    uint32_t uniform(uint32_t v) const { return v; }
    using Same = cudec_amd::zstd::PerLane; /* one lane: its values are the wave's */
    /* a batch's sequence records, all held by the one lane */
    struct Records {
        cudec_amd::zstd::Rec r[cudec_amd::zstd::kBatch];
        void put(uint32_t, uint32_t k, const cudec_amd::zstd::Rec& v) { r[k] = v; }
        cudec_amd::zstd::Rec get(uint32_t k) const { return r[k]; }
        cudec_amd::zstd::Rec own(uint32_t, uint32_t k) const { return r[k]; }
    };
};

}  // namespace

extern "C" {

/* Decodes one frame; on CUDEC_OK *written is the output size and
 * *checksum / *has_checksum what the frame's content checksum must be. */
int32_t host_zstd_decode(const uint8_t* src, uint64_t size, uint8_t* dst, uint64_t cap,
                         uint64_t* written, uint32_t* checksum, int32_t* has_checksum) {
    using namespace cudec_amd::zstd;
    FrameTables* t = static_cast<FrameTables*>(std::calloc(1, sizeof(FrameTables)));
    uint8_t* lits = static_cast<uint8_t*>(std::malloc(kBlockCeiling));
    FrameEnd end{0, 0, false};
    const int st = decode_frame(OneLane{}, src, size, dst, cap, *t, lits, end);
    std::free(lits);
    std::free(t);
    *written = st == CUDEC_OK ? end.written : 0;
    *checksum = end.checksum;
    *has_checksum = end.has_checksum ? 1 : 0;
    return st;
}

/* The code tables, for the closed-form test of their baselines. */
void host_zstd_tables(uint32_t* ll_base, uint8_t* ll_bits, uint32_t* ml_base, uint8_t* ml_bits) {
    using namespace cudec_amd::zstd;
    std::memcpy(ll_base, kLitLenBase, sizeof(kLitLenBase));
    std::memcpy(ll_bits, kLitLenBits, sizeof(kLitLenBits));
    std::memcpy(ml_base, kMatchLenBase, sizeof(kMatchLenBase));
    std::memcpy(ml_bits, kMatchLenBits, sizeof(kMatchLenBits));
}

}  // extern "C"
