/*
 * gputools_amd.h -- C-ABI of the MI355X (gfx950) pairwise-distance library,
 * of the hierarchical-clustering stage built on its output, and of gpuCor's
 * all-pairs correlations (Pearson: the same contraction on standardised
 * variables; Kendall: popcounts over sample-pair bit planes).
 *
 * Two groups of entry points:
 *
 *  1. Drop-in .C exports with the exact names and signatures that gputools'
 *     R code binds through useDynLib(gputools) (NAMESPACE:1) and .C(...).
 *     Every argument is a pointer to R-owned host memory; nothing is kept
 *     after return.  Errors are raised as R errors (Rf_error) when the library
 *     is loaded into an R process, and otherwise recorded for gt_last_error().
 *
 *  2. The native gt_* API: integer status return (GT_OK == 0), 64-bit sizes,
 *     explicit leading dimensions (in elements), device pointers and an
 *     optional HIP stream (NULL = the legacy default stream).  It never throws
 *     or longjmps across the boundary.  Calls on distinct streams are
 *     thread-safe.
 *
 * Layouts
 *   X       row-major n x k FP32, row stride ldx >= k elements (the pitched
 *           layout of src/distance.h:10-19 with the pitch in elements).
 *   PACKED  n(n-1)/2 FP32, R 'dist' order: pair (i, j), i > j, 0-based, at
 *           j*n - j*(j+1)/2 + (i - j - 1)  (what as.dist() keeps at
 *           R/gpuHclust.R:29).
 *   FULL    symmetric n x n FP32, row stride ldd >= n, zero diagonal
 *           (what distance_device fills, src/distance.cpp:209-257).
 *   CROSS   nb x na FP32, row stride ldd >= na: pair (x in A, y in B) at
 *           y*ldd + x (the layout of the *_kernel cross kernels,
 *           inst/cuda/distance.cu:3-38).
 */
#ifndef GPUTOOLS_AMD_H
#define GPUTOOLS_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Distance methods: same values as dist_method (src/distance.h:5-8). */
enum gt_method {
    GT_EUCLIDEAN = 0,
    GT_MAXIMUM = 1,
    GT_MANHATTAN = 2,
    GT_CANBERRA = 3,
    GT_BINARY = 4,
    GT_MINKOWSKI = 5
};

/* Arithmetic of the euclidean path.  FP32 = exact differences on the VALU
 * (every method).  The split forms compute the Gram form
 * ||x||^2 + ||y||^2 - 2 x.y on MFMA with FP32 accumulation and recompute
 * near-coincident pairs (d^2 < tau (||x||^2 + ||y||^2)) from exact FP32
 * differences; euclidean only, max relative error <= 1e-5:
 *   BF16X3  each value split into three BF16 terms, six MFMA products
 *   FP16X3  each row scaled by a power of two, each value split into two
 *           FP16 terms (hi + lo), three MFMA products */
enum gt_precision {
    GT_PREC_FP32 = 0,
    GT_PREC_BF16X3 = 1,
    GT_PREC_FP16X3 = 2
};

enum gt_layout {
    GT_LAYOUT_PACKED = 0,
    GT_LAYOUT_FULL = 1
};

/* Clustering methods: same values as hc_method (src/hcluster.h:5-16). */
enum gt_hc_method {
    GT_HC_SINGLE = 0,
    GT_HC_COMPLETE = 1,
    GT_HC_WPGMA = 2,
    GT_HC_AVERAGE = 3,
    GT_HC_MEDIAN = 4,
    GT_HC_CENTROID = 5,
    GT_HC_FLEXIBLE_GROUP = 6,
    GT_HC_FLEXIBLE = 7,
    GT_HC_WARD = 8,
    GT_HC_MCQUITTY = 9
};

/* Sample handling of the Pearson correlation: same values as whichObs
 * (src/rinterface.cpp:31-50). */
enum gt_use {
    GT_USE_EVERYTHING = 0,   /* every sample; NaN propagates */
    GT_USE_PAIRWISE = 1      /* per pair, the samples where neither value is NaN */
};

enum gt_status {
    GT_OK = 0,
    GT_ERR_INVALID = 1,     /* bad argument: method, size, pointer, p */
    GT_ERR_HIP = 2,         /* a HIP runtime call failed */
    GT_ERR_ALLOC = 3,       /* device or host allocation failed */
    GT_ERR_UNSUPPORTED = 4, /* valid request this build does not implement */
    GT_ERR_WORKSPACE = 5    /* caller workspace too small */
};

/* ------------------------------------------------------------------------ */
/* 1. Drop-in .C exports                                                     */
/* ------------------------------------------------------------------------ */

/* Replaces Rdistances (src/rinterface.cpp:206-214, decl src/rinterface.h:53-55).
 * Fills the full symmetric numPoints x numPoints matrix 'distances' (host),
 * as distance() does (src/distance.cpp:259-316).  Unknown method names select
 * euclidean, as getDistEnum does (src/rinterface.cpp:147-156). */
void Rdistances(const float *points, const int *numPoints, const int *dim,
                float *distances, const char **method, const float *p);

/* Packed variant of Rdistances for R's .C: writes the n(n-1)/2 values of the
 * 'dist' object (lower triangle, column order) into 'distances' directly, so
 * the R side needs neither single(n * n) (which overflows R's integers for
 * n >= 46341, R/gpuHclust.R:25) nor as.dist (R/gpuHclust.R:29).  No
 * counterpart in the reference; same argument conventions as Rdistances. */
void Rdistances_packed(const float *points, const int *numPoints, const int *dim,
                       float *distances, const char **method, const float *p);

/* Rdistances_packed with a choice of arithmetic for R callers: *precision is
 * "fp32" (the FP32 difference kernels, what Rdistances computes), "bf16x3"
 * or "fp16x3" (the split-precision Gram form, euclidean only, max relative
 * error <= 1e-5 against FP64).  No counterpart in the reference, whose
 * gpuDist (R/gpuHclust.R:21-27) has one FP32 path; an unchanged gpuDist
 * keeps calling Rdistances.  Unknown precision names are an error. */
void Rdistances_packed_prec(const float *points, const int *numPoints, const int *dim,
                            float *distances, const char **method, const float *p,
                            const char **precision);

/* Replaces Rdistclust (src/rinterface.cpp:172-204, decl src/rinterface.h:50-52).
 * Distance stage in the device-full layout with p = 2 (src/rinterface.cpp:182),
 * then the clustering stage on that matrix (gt_hclust, lambda = beta = 0.5),
 * then R's hclust merge matrix (column-major (n-1) x 2), order and heights, as
 * formatClustering builds them (src/rinterface.cpp:241-292).  Method names map
 * as getClusterEnum does (gt_hclust_method_from_name). */
void Rdistclust(const char **distmethod, const char **clustmethod,
                const float *points, const int *numPoints, const int *dim,
                int *merge, int *order, float *val);

/* Replaces Rhcluster (src/rinterface.cpp:216-239, decl src/rinterface.h:56-57):
 * clustering of a host numPoints x numPoints distance matrix (what gpuHclust
 * passes, R/gpuHclust.R:70-77), lambda = beta = 0.5, R-formatted results. */
void Rhcluster(const float *distMat, const int *numPoints, int *merge, int *order, float *val,
               const char **method);

/* Replaces rpmcc (src/rinterface.cpp:33-55, body pmcc at
 * src/correlation.cpp:66-160), what gpuCor(method = "pearson") calls
 * (R/gpuCor.R:33-37).  Variable x of A is the sampleSize floats at
 * samplesA + x * sampleSize (a column of R's matrix).  Writes
 * numSamplesA x numSamplesB row-major numPairs, correlations and t-scores
 * (hostSignif, src/correlation.cpp:51-64).  whichObs 0 = "everything"
 * (split-precision MFMA, FP32-equivalent), 1 = "pairwise.complete.obs"
 * (FP32).  Identical A and B contents are computed as one symmetric matrix. */
void rpmcc(const int *whichObs, const float *samplesA, const int *numSamplesA,
           const float *samplesB, const int *numSamplesB, const int *sampleSize, float *numPairs,
           float *correlations, float *signifs);

/* Replaces rtestT (src/rinterface.cpp:87-91, body testSignif at
 * src/correlation.cpp:17-49), what gpuTtest calls (R/gpuCor.R:68-77):
 * t = coeff * sqrt((pairs - 2) / (1 - coeff^2)) for n entries, on the GPU,
 * without hostSignif's cap. */
void rtestT(const float *pairs, const float *coeffs, const int *n, float *ts);

/* Replaces RgpuKendall (src/rinterface.cpp:128-132, body masterKendall at
 * src/kendall.cpp:13-55, kernel inst/cuda/kendall.cu:4-47), what
 * gpuCor(method = "kendall") calls (R/gpuCor.R:52-55).  Variable x of X is
 * the sampleSize floats at X + x * sampleSize.  answers: nx x ny column-major
 * (pair (x, y) at y * nx + x), Kendall's tau as the reference defines it
 * (sample pairs tied in both variables count as concordant). */
void RgpuKendall(const float *X, const int *nx, const float *Y, const int *ny,
                 const int *sampleSize, double *answers);

/* Replace cuCompile / unloadPackage (src/cudaUtils.cpp:117-165, 167-195),
 * called from R/zzz.R:34-38 and :41-43.  Kernels are precompiled gfx950 code
 * objects, so both only record that the package was loaded/unloaded. */
void cuCompile(const int *numFiles, const char **cuFilenames, const char **cuSrc);
void unloadPackage(void);

/* Replace rsetDevice / rgetDevice (src/rinterface.cpp:79-85, bodies
 * src/correlation.cpp:163-175): range-checked hipSetDevice / hipGetDevice. */
void rsetDevice(const int *device);
void rgetDevice(int *device);

/* ------------------------------------------------------------------------ */
/* 2. Native API                                                             */
/* ------------------------------------------------------------------------ */

/* Message of the last failed call on this thread ("" if none). */
const char *gt_last_error(void);

/* Library version, major*10000 + minor*100 + patch. */
int gt_version(void);

/* getDistEnum (src/rinterface.cpp:147-156): exact name -> method, anything
 * else -> GT_EUCLIDEAN. */
int gt_method_from_name(const char *name);

/* Bytes of device workspace gt_dist_ws needs for this request (0 if none). */
int gt_dist_workspace_size(int64_t n, int64_t k, int method, int precision,
                           size_t *bytes);

/* Device-resident all-pairs distances of the rows of X.  The replacement of
 * distance_device (src/distance.cpp:209-257) with a packed or full output.
 * Workspace, when the request needs one, is taken from a stream-ordered pool
 * (hipMallocAsync / hipFreeAsync on 'stream'). */
int gt_dist(const float *x, int64_t n, int64_t k, int64_t ldx, int method, float p,
            int precision, int layout, float *out, int64_t ldd, void *stream);

/* Same, with a caller-provided device workspace of gt_dist_workspace_size
 * bytes (no allocation inside). */
int gt_dist_ws(const float *x, int64_t n, int64_t k, int64_t ldx, int method, float p,
               int precision, int layout, float *out, int64_t ldd,
               void *workspace, size_t workspace_bytes, void *stream);

/* Bytes of the static device scratch behind gt_dist_noalloc (per GPU). */
#define GT_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)

/* gt_dist without any allocation: no counterpart in the reference.  For
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  The workspace
 * gt_dist_workspace_size reports comes from a static device array of
 * GT_NOALLOC_SCRATCH_BYTES that every GPU the library runs on holds until the
 * process ends; a request that needs more returns GT_ERR_WORKSPACE (use
 * gt_dist_ws).  The plain FP32 methods other than binary need no workspace;
 * binary needs n * (ceil(k/32) + 1) * 4 bytes, the split-precision forms
 * about n * k * 6 (BF16X3) or n * k * 4 (FP16X3) bytes.  Every call on one
 * GPU shares that GPU's array: calls that overlap on one GPU (two streams,
 * two host threads) race and must be serialised by the caller.  Results are
 * those of gt_dist, bit for bit. */
int gt_dist_noalloc(const float *x, int64_t n, int64_t k, int64_t ldx, int method, float p,
                    int precision, int layout, float *out, int64_t ldd, void *stream);

/* Cross distances between the rows of A (na x k) and B (nb x k), CROSS
 * layout.  The semantics of the unreachable *_kernel cross kernels
 * (inst/cuda/distance.cu:3-38, 80-114, 156-190, 234-271, 318-364, 417-451),
 * FP32 path, with R's canberra rule. */
int gt_dist_cross(const float *a, int64_t na, int64_t lda, const float *b, int64_t nb,
                  int64_t ldb, int64_t k, int method, float p, float *out, int64_t ldd,
                  void *stream);

/* Host-to-host call: the counterpart of distance() (src/distance.cpp:259-316)
 * with either output layout; synchronous. */
int gt_dist_host(const float *x_host, int64_t n, int64_t k, int method, float p,
                 int precision, int layout, float *out_host);

/* distanceLeaveOnGpu (src/distance.cpp:318-344): copies the host points to the
 * device, computes the FULL layout into a new device buffer whose ownership
 * passes to the caller (*d_out, row pitch in bytes in *pitch_bytes); release
 * it with gt_free. */
int gt_dist_device_full(const float *x_host, int64_t n, int64_t k, int method, float p,
                        float **d_out, size_t *pitch_bytes);

int gt_free(void *device_ptr);

/* getClusterEnum (src/rinterface.cpp:158-170): exact names "complete",
 * "wpgma", "average", "median", "centroid", "flexible_group", "flexible",
 * "ward", "mcquitty"; anything else (including "flexible group") ->
 * GT_HC_SINGLE. */
int gt_hclust_method_from_name(const char *name);

/* Bytes of device workspace gt_hclust needs for n points. */
int gt_hclust_workspace_size(int64_t n, size_t *bytes);

/* Hierarchical clustering of the full, symmetric n x n FP32 device matrix d
 * (row stride ldd elements), which is destroyed: the replacement of
 * hclusterPreparedDistances (src/hcluster.cpp:40-201) and its kernels
 * (inst/cuda/hcluster.cu:4-400).  Writes the n - 1 merges to device arrays:
 * at step s the clusters represented by sub[s] < sup[s] (0-based point
 * indices) merge at height[s]; the merged cluster keeps index sup[s].  Ties
 * and arithmetic follow the reference bit for bit.  n below 2^24. */
int gt_hclust(float *d, int64_t n, int64_t ldd, int method, float lambda, float beta, int *sub,
              int *sup, float *height, void *workspace, size_t workspace_bytes, void *stream);

/* Where the clustering kernel keeps its per-row state (counts, cached row
 * minima, the rewritten row): GT_HC_STATE_AUTO as gt_hclust does (LDS up to
 * GT_HC_LDS_MAX_N points, global memory above), or one placement forced (LDS
 * only up to GT_HC_LDS_MAX_N).  Both placements run the same merge rounds and
 * give the same merges and heights bit for bit. */
enum gt_hc_state {
    GT_HC_STATE_AUTO = 0,
    GT_HC_STATE_LDS = 1,
    GT_HC_STATE_GLOBAL = 2
};
#define GT_HC_LDS_MAX_N 9216

/* gt_hclust with the per-row state placement chosen by the caller, so that
 * either kernel variant can be run (and compared) at any size.  The
 * reference has one placement (its kernels re-read the whole matrix every
 * round, src/hcluster.cpp:40-201); this entry point has no counterpart. */
int gt_hclust_state(float *d, int64_t n, int64_t ldd, int method, float lambda, float beta,
                    int *sub, int *sup, float *height, void *workspace, size_t workspace_bytes,
                    int state, void *stream);

/* gt_hclust on a host matrix (the replacement of hcluster,
 * src/hcluster.cpp:15-38): host in, host sub/sup/height out, synchronous. */
int gt_hclust_host(const float *d_host, int64_t n, int64_t ldd, int method, float lambda,
                   float beta, int *sub, int *sup, float *height);

/* R's hclust merge matrix (merge[s] and merge[s + len], 1-based, negative =
 * singleton) and order (len + 1 entries, 1-based) from a merge sequence of
 * length len: formatClustering / getPrintOrder (src/rinterface.cpp:241-292). */
int gt_hclust_format(int64_t len, const int *sub, const int *sup, int *merge, int *order);

/* Backward of the packed distance vector: dX (n x k, row stride lddx) =
 * sum_j G_ij * d(d_ij)/d(x_i) for euclidean (any precision flag is ignored:
 * FP32) and minkowski (p >= 1).  G is packed dL/dD.  D is recomputed
 * tile-by-tile and never stored.  Pairs with d_ij == 0 contribute 0. */
int gt_dist_backward(const float *x, int64_t n, int64_t k, int64_t ldx, int method,
                     float p, const float *g_packed, float *dx, int64_t lddx,
                     void *stream);

/* Fused metric-MDS stress S = sum_{i<j} (||y_i - y_j|| - delta_ij)^2 with
 * delta packed.  Writes dS/dY (n x q, stride lddy) when dy != NULL and the
 * stress into *stress_dev (one double on the device; FP64 is used only for
 * this scalar reduction).  D is recomputed per tile, never stored. */
int gt_mds_stress(const float *y, int64_t n, int64_t q, int64_t ldy,
                  const float *delta_packed, float *dy, int64_t lddy,
                  double *stress_dev, void *stream);

/* Batched distances over many independent point sets in one persistent
 * launch (no counterpart in the reference, which computes one matrix per
 * distance() call, src/distance.cpp:259-316).  Set s has sizes[s] rows of
 * dims[s] features starting at element x_offsets[s] of x (row stride
 * dims[s]); its packed output ('dist' order) starts at element out_offsets[s]
 * of out.  Methods: euclidean, manhattan, maximum (FP32 path).
 *
 * gt_dist_batched_schedule builds, on the host, the tile schedule for the
 * given host copies of sizes and dims and a tile shape (1..gt_batched_tile_count():
 * 32, 64 or 128 pairs per tile side): *count receives the number of int64
 * entries; they are written to schedule_host when capacity allows.
 * gt_dist_batched takes the descriptor arrays and that schedule in device
 * memory, allocates nothing and synchronises nothing, so it can be captured
 * in a HIP graph and replayed.  Entries that do not describe a tile of a set
 * are skipped. */
int gt_batched_tile_count(void);
int gt_dist_batched_schedule(const int64_t *sizes_host, const int64_t *dims_host, int64_t nsets,
                             int tile, int64_t *schedule_host, int64_t capacity, int64_t *count);
int gt_dist_batched(const float *x, const int64_t *x_offsets, const int64_t *sizes,
                    const int64_t *dims, int64_t nsets, float *out, const int64_t *out_offsets,
                    int method, int tile, const int64_t *schedule, int64_t count, void *stream);

/* Pearson correlation of every variable of A (na x dim, row stride lda >= dim)
 * with every variable of B (nb x dim, stride ldb), the device form of pmcc
 * (src/correlation.cpp:66-160).  Outputs na x nb, row stride ldo >= nb, pair
 * (x, y) at x * ldo + y: the sample count, the correlation and hostSignif's
 * t-score.  b == NULL (or b == a with the same shape) correlates A with
 * itself: the result is exactly symmetric and half the tiles are computed.
 * use GT_USE_EVERYTHING: precision FP32 (VALU), BF16X3 or FP16X3 (MFMA on
 * standardised variables); GT_USE_PAIRWISE: FP32 only.  The workspace
 * (gt_pmcc_workspace_size, same != 0 for the self case) holds the
 * standardised variables; none is needed for GT_USE_PAIRWISE. */
int gt_pmcc_workspace_size(int use, int precision, int64_t na, int64_t nb, int64_t dim, int same,
                           size_t *bytes);
int gt_pmcc(int use, int precision, const float *a, int64_t na, int64_t lda, const float *b,
            int64_t nb, int64_t ldb, int64_t dim, float *pairs, float *corr, float *ts, int64_t ldo,
            void *workspace, size_t workspace_bytes, void *stream);
/* Host-memory form (contiguous rows; b_host == NULL for the self case). */
int gt_pmcc_host(int use, int precision, const float *a_host, int64_t na, const float *b_host,
                 int64_t nb, int64_t dim, float *pairs, float *corr, float *ts);
/* gpuSignif (inst/cuda/correlation.cu:292-310) on device arrays of n entries. */
int gt_ttest(const float *pairs, const float *corr, int64_t n, float *ts, void *stream);

/* Kendall's tau of every variable of A (na x n, row stride lda >= n) with
 * every variable of B (nb x n), the device form of masterKendall
 * (src/kendall.cpp:13-55): tau (FP64) nb x na, row stride ldt >= na, pair
 * (x, y) at y * ldt + x (the reference's layout).  Concordant sample pairs
 * are counted exactly in integers; the quotient follows the reference's FP32
 * / FP64 steps (inst/cuda/kendall.cu:40-45).  b == NULL (or b == a with the
 * same shape) is the self case.  The workspace (gt_kendall_workspace_size)
 * holds three bit planes of n(n-1)/2 bits per variable, per-variable flags,
 * a small per-chunk table of the bits past row ends, and the counts. */
int gt_kendall_workspace_size(int64_t na, int64_t nb, int64_t n, int same, size_t *bytes);
int gt_kendall(const float *a, int64_t na, int64_t lda, const float *b, int64_t nb, int64_t ldb,
               int64_t n, double *tau, int64_t ldt, void *workspace, size_t workspace_bytes,
               void *stream);
/* Host-memory form (contiguous rows; b_host == NULL for the self case). */
int gt_kendall_host(const float *a_host, int64_t na, const float *b_host, int64_t nb, int64_t n,
                    double *tau);

#ifdef __cplusplus
}
#endif
#endif /* GPUTOOLS_AMD_H */
