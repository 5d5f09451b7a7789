// nbx_mixed.hip -- precision form (FP32 node-relative search with an FP64
// re-check band).  Launcher stub until the kernel lands.
#include "nbx_kernels.h"

hipError_t nbx_launch_knn_mixed(const DevTree &T, const KnnArgs &A, double band, int32_t *fallback_list,
                                int32_t *fallback_count, hipStream_t st) {
    (void)T; (void)A; (void)band; (void)fallback_list; (void)fallback_count; (void)st;
    return hipErrorNotSupported;
}
