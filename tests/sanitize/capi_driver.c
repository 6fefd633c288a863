/* Calls the C-ABI's argument-checking paths (no GPU is touched: every call
 * is rejected before the device) in a build of the C-ABI layer with
 * AddressSanitizer and UndefinedBehaviorSanitizer on its host code. */
#include <stdio.h>
#include <string.h>

#include "../../include/gputools_amd.h"

#define EXPECT(call, status)                                                        \
    do {                                                                            \
        const int s_ = (call);                                                      \
        if (s_ != (status)) {                                                       \
            fprintf(stderr, "line %d: %s returned %d (%s)\n", __LINE__, #call, s_,  \
                    gt_last_error());                                               \
            return 1;                                                               \
        }                                                                           \
    } while (0)

int main(void)
{
    float *dummy = (float *)16;
    size_t bytes = 0;
    EXPECT(gt_dist(dummy, 10, 4, 4, 9, 2.0f, 0, 0, dummy, 10, NULL), GT_ERR_INVALID);
    EXPECT(gt_dist(dummy, -1, 4, 4, 0, 2.0f, 0, 0, dummy, 10, NULL), GT_ERR_INVALID);
    EXPECT(gt_dist(dummy, 10, 4, 3, 0, 2.0f, 0, 0, dummy, 10, NULL), GT_ERR_INVALID);
    EXPECT(gt_dist(dummy, 10, 4, 4, GT_MINKOWSKI, -1.0f, 0, 0, dummy, 10, NULL), GT_ERR_INVALID);
    EXPECT(gt_dist(dummy, 10, 4, 4, GT_MANHATTAN, 2.0f, GT_PREC_BF16X3, 0, dummy, 10, NULL),
           GT_ERR_UNSUPPORTED);
    EXPECT(gt_dist(dummy, 10, 4, 4, 0, 2.0f, 0, GT_LAYOUT_FULL, dummy, 9, NULL), GT_ERR_INVALID);
    EXPECT(gt_dist_ws(dummy, 100, 64, 64, 0, 2.0f, GT_PREC_BF16X3, 0, dummy, 100, NULL, 0, NULL),
           GT_ERR_WORKSPACE);
    EXPECT(gt_dist_noalloc(dummy, 65536, 512, 512, 0, 2.0f, GT_PREC_BF16X3, 0, dummy, 65536, NULL),
           GT_ERR_WORKSPACE);
    EXPECT(gt_dist_workspace_size(65536, 512, 0, GT_PREC_BF16X3, &bytes), GT_OK);
    if (bytes < (size_t)65536 * 512 * 6) return 1;
    EXPECT(gt_dist_workspace_size(-1, 2, 0, 0, &bytes), GT_ERR_INVALID);
    EXPECT(gt_dist_cross(dummy, 5, 3, dummy, 4, 2, 3, 0, 2.0f, dummy, 5, NULL), GT_ERR_INVALID);
    EXPECT(gt_hclust(dummy, 10, 9, 0, 0.5f, 0.5f, NULL, NULL, NULL, NULL, 0, NULL), GT_ERR_INVALID);
    EXPECT(gt_hclust_state(dummy, 9300, 9300, 0, 0.5f, 0.5f, (int *)dummy, (int *)dummy, dummy,
                           dummy, 1 << 20, GT_HC_STATE_LDS, NULL),
           GT_ERR_INVALID);
    EXPECT(gt_hclust_state(dummy, 10, 10, 0, 0.5f, 0.5f, (int *)dummy, (int *)dummy, dummy, dummy,
                           1 << 20, 7, NULL),
           GT_ERR_INVALID);
    int sub[3] = {0, 1, 2}, sup[3] = {1, 2, 3}, merge[6], order[4];
    EXPECT(gt_hclust_format(3, sub, sup, merge, order), GT_OK);
    if (merge[0] != -1 || merge[3] != -2 || merge[2] != -4 || merge[5] != 2) return 1;
    int bad_sup[3] = {1, 1, 3};
    EXPECT(gt_hclust_format(3, sub, bad_sup, merge, order), GT_ERR_INVALID);
    EXPECT(gt_pmcc(7, 0, dummy, 4, 4, NULL, 4, 4, 4, dummy, dummy, dummy, 4, NULL, 0, NULL),
           GT_ERR_INVALID);
    EXPECT(gt_kendall(dummy, 4, 3, NULL, 4, 3, 4, (double *)dummy, 4, NULL, 0, NULL),
           GT_ERR_INVALID);
    EXPECT(gt_dist_backward(dummy, 10, 4, 4, GT_CANBERRA, 2.0f, dummy, dummy, 4, NULL),
           GT_ERR_UNSUPPORTED);
    EXPECT(gt_mds_stress(dummy, 10, 0, 1, dummy, NULL, 1, NULL, NULL), GT_ERR_INVALID);
    int64_t count = 0;
    int64_t sizes[2] = {5, 7}, dims[2] = {3, 0};
    EXPECT(gt_dist_batched_schedule(sizes, dims, 2, 1, NULL, 0, &count), GT_ERR_INVALID);
    if (gt_method_from_name("canberra") != GT_CANBERRA || gt_method_from_name("x") != GT_EUCLIDEAN)
        return 1;
    if (gt_hclust_method_from_name("flexible group") != GT_HC_SINGLE) return 1;
    /* the .C exports record their errors (no R in this process) */
    const int n = 3, dim = 2;
    const char *method = "euclidean", *prec = "fp8";
    const float p = 2.0f;
    float pts[6] = {0}, out[3];
    Rdistances_packed_prec(pts, &n, &dim, out, &method, &p, &prec);
    if (strstr(gt_last_error(), "unknown precision") == NULL) return 1;
    const int w = 5;
    rpmcc(&w, pts, &n, pts, &n, &dim, out, out, out);
    if (strstr(gt_last_error(), "unknown use") == NULL) return 1;
    printf("C-ABI sanitizer run: ok\n");
    return 0;
}
