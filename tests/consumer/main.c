/* Links cudec::cudec from an installed prefix and calls through the C ABI
 * without a device: the version, the result layout, and the synchronous
 * argument rejects, which make no device call. Prints "consumer ok". */
#include <stddef.h>
#include <stdio.h>

#include "cudec.h"
#include "cudec_mi355x.h"
#include "cudec_snappy.h"

int main(void) {
    if (cudec_version() != 100) {
        printf("cudec_version %d\n", cudec_version());
        return 1;
    }
    if (sizeof(cudec_chunk_result) != 16 || offsetof(cudec_chunk_result, bytes_written) != 8) {
        printf("cudec_chunk_result layout\n");
        return 1;
    }
    if (cudec_lz4_decompress_batch(NULL, NULL, NULL, NULL, 1, NULL, NULL) !=
            CUDEC_ERR_INVALID_ARGUMENT ||
        cudec_snappy_decompress_batch(NULL, NULL, NULL, NULL, 1, NULL, NULL) !=
            CUDEC_ERR_INVALID_ARGUMENT ||
        cudec_mi355x_max_batch_chunks() == 0) {
        printf("argument rejects\n");
        return 1;
    }
    size_t written = 7;
    if (cudec_lz4f_decompress(NULL, 0, NULL, 0, &written) != CUDEC_ERR_INVALID_ARGUMENT ||
        written != 0) {
        printf("frame rejects\n");
        return 1;
    }
    printf("consumer ok\n");
    return 0;
}
