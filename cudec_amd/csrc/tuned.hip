/* The tuned launch form. */
#include "common.h"

namespace cudec_amd {

hipError_t launch_tuned(int format, const void* const* srcs,
                        const size_t* src_sizes, void* const* dsts,
                        const size_t* caps, size_t n,
                        cudec_chunk_result* results, hipStream_t stream) {
    return launch_plain(format, srcs, src_sizes, dsts, caps, n, results,
                        stream);
}

}  // namespace cudec_amd
