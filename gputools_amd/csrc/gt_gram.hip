// gt_gram.hip -- split-precision MFMA Gram path for the euclidean distance.
#include "gt_common.h"
#include "gt_internal.h"

namespace gt {

size_t gram_workspace_bytes(int64_t n, int64_t k)
{
    (void)k;
    return ((size_t)n * sizeof(float) + 255) / 256 * 256;
}

hipError_t launch_gram(int precision, int layout, const float *x, int64_t n, int64_t k,
                       int64_t ldx, float *out, int64_t ldd, void *workspace, hipStream_t stream)
{
    (void)precision; (void)layout; (void)x; (void)n; (void)k; (void)ldx; (void)out; (void)ldd;
    (void)workspace; (void)stream;
    return hipErrorNotSupported;
}

}  // namespace gt
