/* Driver over iderex/cudec's own Zstd frame twin, compiled from the
 * reference tree where it lies (see oracle/Makefile, target `ref`); the
 * output goes to oracle/_ref/ and is never committed. TEST INFRASTRUCTURE
 * ONLY: it is the checker of the device Zstd decoder's verdicts.
 *
 * The reference's tests/zstd_twin_driver.h (cudec_twin::DecodeFrame) runs
 * the shipped units src/zstd_frame.h, zstd_literals.h, zstd_seq.h,
 * zstd_repcode.h and xxhash64.h over one whole frame; it is included
 * unmodified. What is written here is only the C entry that copies its
 * output out. */
#include "zstd_twin_driver.h"

#include <cstdint>
#include <cstring>

extern "C" {

/* One frame: status, and on CUDEC_OK the decoded bytes in dst[0, *written).
 * *stage receives the stage that refused (cudec_twin::DriverStage). */
int32_t ref_zstd_decode(const uint8_t* src, uint64_t src_size, uint8_t* dst, uint64_t cap,
                        uint64_t* written, int32_t* stage) {
    cudec_twin::Bytes frame(src, src + src_size);
    const cudec_twin::Run run = cudec_twin::DecodeFrame(frame, cap);
    *stage = static_cast<int32_t>(run.stage);
    if (!run.ok) {
        *written = 0;
        return static_cast<int32_t>(run.status);
    }
    if (!run.output.empty()) {
        std::memcpy(dst, run.output.data(), run.output.size());
    }
    *written = run.output.size();
    return CUDEC_OK;
}

}  // extern "C"
