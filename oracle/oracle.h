/*
 * oracle/oracle.h -- CPU restatement of gputools' pairwise-distance path
 * (R stats::dist semantics, FP64), of its hierarchical-clustering stage
 * (the reference's FP32 merge loop) and of gpuCor's correlations.  TEST
 * INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py, never by the product library.
 */
#ifndef GPUTOOLS_ORACLE_H
#define GPUTOOLS_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Same values as dist_method (src/distance.h:5-8). */
enum {
    ORACLE_EUCLIDEAN = 0, ORACLE_MAXIMUM = 1, ORACLE_MANHATTAN = 2,
    ORACLE_CANBERRA = 3, ORACLE_BINARY = 4, ORACLE_MINKOWSKI = 5
};

int64_t oracle_packed_index(int64_t n, int64_t i, int64_t j);
double oracle_pair(const float *a, const float *b, int64_t k, int method, double p);
int oracle_dist_packed(const float *x, int64_t n, int64_t k, int64_t ldx,
                       int method, double p, double *out, int nthreads);
int oracle_dist_pairs(const float *x, int64_t k, int64_t ldx,
                      const int64_t *ii, const int64_t *jj, int64_t npairs,
                      int method, double p, double *out, int nthreads);
int oracle_dist_cross(const float *a, int64_t na, const float *b, int64_t nb,
                      int64_t k, int method, double p, double *out, int nthreads);
double oracle_compare_packed(const float *x, int64_t n, int64_t k, int64_t ldx,
                             int method, double p, const float *got,
                             double abs_floor, int64_t *worst_i, int64_t *worst_j,
                             int nthreads);
int oracle_dist_backward(const float *x, int64_t n, int64_t k, int method, double p,
                         const double *g, double *dx, int nthreads);
/* Same values as hc_method (src/hcluster.h:5-16). */
enum {
    ORACLE_HC_SINGLE = 0, ORACLE_HC_COMPLETE = 1, ORACLE_HC_WPGMA = 2, ORACLE_HC_AVERAGE = 3,
    ORACLE_HC_MEDIAN = 4, ORACLE_HC_CENTROID = 5, ORACLE_HC_FLEXIBLE_GROUP = 6,
    ORACLE_HC_FLEXIBLE = 7, ORACLE_HC_WARD = 8, ORACLE_HC_MCQUITTY = 9
};

/* hclust_oracle.c: d is a full n x n FP32 matrix (row stride ldd), destroyed. */
int oracle_hclust(float *d, int64_t n, int64_t ldd, int method, float lambda, float beta,
                  int *sub, int *sup, float *height);
int oracle_hclust_format(int len, const int *sub, const int *sup, int *merge, int *order);

/* cor_oracle.c: Pearson correlation (pmcc) in FP64, and the FP32 t-scores. */
int oracle_pmcc(int use, const float *a, int64_t na, const float *b, int64_t nb, int64_t dim,
                double *pairs, double *corr, int nthreads);
void oracle_signif(const float *pairs, const float *corr, int64_t n, int cap, float *ts);
/* Kendall's tau as gpuKendall defines it; tau and count nb x na ([y * na + x]). */
int oracle_kendall(const float *a, int64_t na, const float *b, int64_t nb, int64_t n, double *tau,
                   uint64_t *count, int nthreads);

double oracle_mds_stress(const float *y, int64_t n, int64_t q, const double *delta,
                         double *dy, int nthreads);

#ifdef __cplusplus
}
#endif
#endif
