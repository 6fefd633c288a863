// gt_internal.h -- host-side declarations shared by the library's translation
// units (kernel launchers and the C-ABI layer).  Not part of the public ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace gt {

// Per-pair accumulation of the FP32 VALU path.
enum PlainOp {
    OP_EUCLID = 0,     // sum (a-b)^2, then sqrt            (distance.cu:40-78)
    OP_MAXIMUM = 1,    // max |a-b|                          (distance.cu:116-154)
    OP_MANHATTAN = 2,  // sum |a-b|                          (distance.cu:192-232)
    OP_CANBERRA = 3,   // sum |a-b|/|a+b|, R's 0/0 rule       (distance.cu:273-316)
    OP_MINK3 = 4,      // sum |a-b|^3, then cube root        (distance.cu:453-493, p = 3)
    OP_MINKP = 5,      // sum |a-b|^p, then ^(1/p)           (distance.cu:453-493)
    OP_DOT = 6         // sum a*b (Pearson on standardised variables, correlation.cu:252-290)
};

enum OutLayout {
    OUT_PACKED = 0,  // R 'dist' order
    OUT_FULL = 1,    // symmetric n x n, zero diagonal
    OUT_CROSS = 2    // nb x na, pair (x in A, y in B) at y*ldd + x
};

// Plain FP32 tile kernels.  Self mode: a == b, na == nb == n, tiles on and
// below the diagonal.  Cross mode: every (A tile, B tile).
hipError_t launch_plain(int op, int layout, const float *a, int64_t na, int64_t lda,
                        const float *b, int64_t nb, int64_t ldb, int64_t k, float p,
                        float *out, int64_t ldd, hipStream_t stream);

// Binary method: rows -> bit rows (x != 0), then popcount tiles.
int64_t binary_words(int64_t k);
int64_t binary_ldw(int64_t k);   // packed row stride: the words and the row's bit count
hipError_t launch_binary_pack(const float *x, int64_t n, int64_t k, int64_t ldx,
                              uint32_t *bits, int64_t ldw, hipStream_t stream);
hipError_t launch_binary(int layout, const uint32_t *a, int64_t na, const uint32_t *b,
                         int64_t nb, int64_t words, int64_t ldw, float *out, int64_t ldd,
                         hipStream_t stream);
// One point set (the self distances): packed layouts below 2^24 points count
// on the matrix cores (one byte per feature), the others on the bit rows.
size_t binary_workspace_bytes(int64_t n, int64_t k);
hipError_t launch_binary_self(int layout, const float *x, int64_t n, int64_t k, int64_t ldx,
                              void *workspace, float *out, int64_t ldd, hipStream_t stream);

// Fill n floats with a value (used for the k == 0 case: every distance NA).
hipError_t launch_fill(float *out, int64_t count, float value, hipStream_t stream);
hipError_t launch_fill_full(float *out, int64_t n, int64_t ldd, float offdiag,
                            hipStream_t stream);

// Split-precision MFMA Gram path (euclidean).  Workspace: split terms, row
// norms and row scales (gram_workspace_bytes).
size_t gram_workspace_bytes(int64_t n, int64_t k, int precision);
hipError_t launch_gram(int precision, int layout, const float *x, int64_t n, int64_t k,
                       int64_t ldx, float *out, int64_t ldd, void *workspace,
                       hipStream_t stream);

// Backward (dL/dX) and fused MDS stress.  op: OP_EUCLID, OP_MINK3 or OP_MINKP.
// workspace: backward_workspace_bytes(n, k) (partial dX slices when the j
// blocks are split over workgroups; 0 otherwise).
size_t backward_workspace_bytes(int64_t n, int64_t k);
hipError_t launch_backward(int op, const float *x, int64_t n, int64_t k, int64_t ldx,
                           float p, const float *g_packed, float *dx, int64_t lddx,
                           float *workspace, hipStream_t stream);
// partials: mds_partials_count(n, q) doubles; workspace as for the backward.
int64_t mds_partials_count(int64_t n, int64_t q);
hipError_t launch_mds(const float *y, int64_t n, int64_t q, int64_t ldy,
                      const float *delta, float *dy, int64_t lddy, double *stress,
                      double *partials, float *workspace, hipStream_t stream);

// Batched persistent kernel.  tile: 1..3 (32, 64, 128 pairs per side).
int batched_tile_side(int tile);
int batched_resident_blocks(int op, int tile);
hipError_t launch_batched(int op, int tile, const float *x, const int64_t *x_off,
                          const int64_t *sizes, const int64_t *dims, const int64_t *out_off,
                          float *out, const int64_t *schedule, int64_t count, int64_t nsets,
                          int grid, hipStream_t stream);

// Hierarchical clustering methods, in hc_method order (src/hcluster.h:5-16).
enum HcMethod {
    HC_SINGLE = 0,
    HC_COMPLETE = 1,
    HC_WPGMA = 2,
    HC_AVERAGE = 3,
    HC_MEDIAN = 4,
    HC_CENTROID = 5,
    HC_FLEXIBLE_GROUP = 6,
    HC_FLEXIBLE = 7,
    HC_WARD = 8,
    HC_MCQUITTY = 9
};

// Clustering on a full n x n device matrix d (row stride ldd, destroyed).
// count, minv, minc: n entries of workspace each; sub, sup, height: n - 1.
// state: HC_STATE_AUTO keeps the per-row state in LDS up to HC_LDS_MAX_N
// points and in global memory above; HC_STATE_LDS / HC_STATE_GLOBAL force a
// placement (LDS only up to HC_LDS_MAX_N).
enum HcState { HC_STATE_AUTO = 0, HC_STATE_LDS = 1, HC_STATE_GLOBAL = 2 };
constexpr int64_t HC_LDS_MAX_N = 9216;
hipError_t launch_hclust(float *d, int64_t n, int64_t ldd, int method, float lambda, float beta,
                         float *count, float *minv, int *minc, int *sub, int *sup, float *height,
                         int state, hipStream_t stream);

// Pearson correlation (gpuCor).  use: 0 everything, 1 pairwise.complete.obs.
// precision (use 0 only): 0 FP32 VALU, 1 BF16 three-term, 2 FP16 two-term
// split MFMA.  a: na x dim variables (row stride lda), b likewise; same means
// b is a (symmetric result).  Outputs na x nb, row stride ldo.
size_t pmcc_workspace_bytes(int use, int precision, int64_t na, int64_t nb, bool same,
                            int64_t dim);
hipError_t launch_pmcc(int use, int precision, const float *a, int64_t na, int64_t lda,
                       const float *b, int64_t nb, int64_t ldb, bool same, int64_t dim,
                       float *pairs, float *corr, float *ts, int64_t ldo, void *workspace,
                       hipStream_t stream);
// gpuTtest: t-scores of n (pairs, corr) without hostSignif's cap.
hipError_t launch_ttest(const float *pairs, const float *corr, float *ts, int64_t n,
                        hipStream_t stream);

// Kendall's tau (gpuCor method = "kendall").  a: na x n variables (row
// stride lda), b likewise; same means b is a.  tau: nb x na FP64, pair (x, y)
// at y * ldt + x (the reference's layout).  Workspace: bit planes and counts.
size_t kendall_workspace_bytes(int64_t na, int64_t nb, bool same, int64_t n);
hipError_t launch_kendall(const float *a, int64_t na, int64_t lda, const float *b, int64_t nb,
                          int64_t ldb, bool same, int64_t n, double *tau, int64_t ldt,
                          void *workspace, hipStream_t stream);

}  // namespace gt
