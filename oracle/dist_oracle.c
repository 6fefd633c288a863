/*
 * oracle/dist_oracle.c -- FP64 CPU restatement of the pairwise-distance path.
 *
 * TEST INFRASTRUCTURE ONLY.  The HIP library in gputools_amd/ is tested
 * against this file: only tests/, __graft_entry__.smoke() and the cpu_baseline
 * leg of bench.py may load it.  The product never links it and never falls back
 * to it.
 *
 * What it restates
 * ----------------
 * gputools' gpuDist() (R/gpuHclust.R:1-39) calls .C("Rdistances")
 * (src/rinterface.cpp:206-214) -> distance() (src/distance.cpp:259-316) ->
 * one of the six *_kernel_same kernels of inst/cuda/distance.cu.  The per-pair
 * formulas come from those kernels:
 *   euclidean  distance.cu:40-78    sqrt(sum (a-b)^2)
 *   maximum    distance.cu:116-154  max |a-b|
 *   manhattan  distance.cu:192-232  sum |a-b|
 *   canberra   distance.cu:273-316  sum |a-b|/|a+b|
 *   binary     distance.cu:366-415  #(a xor b) / #(a or b), 0 when #or == 0
 *   minkowski  distance.cu:453-493  (sum |a-b|^p)^(1/p)
 * The oracle the path is judged against is R's stats::dist in FP64, so where
 * the gputools kernel and R's documented semantics (?dist) differ, this file
 * follows R:
 *   - canberra: a term with zero numerator and zero denominator (a == b == 0)
 *     is omitted, and the sum is scaled by k/count; a == -b != 0 gives Inf;
 *     count == 0 gives NA (NaN here).  gputools skips every term whose
 *     denominator is 0 and never rescales (distance.cu:291-296).
 *   - k == 0: every method gives NA (NaN here).
 * Output order is R's 'dist' order (as.dist at R/gpuHclust.R:29): the strict
 * lower triangle in column order, i.e. pair (i, j), i > j, 0-based, at
 *   j*n - j*(j+1)/2 + (i - j - 1).
 * The dead CPU helpers of src/distance.cpp:17-207 are NOT followed (they index
 * the output by the feature count instead of n).
 *
 * Summation is sequential in double, as R does it.  Inputs are the FP32 values
 * the GPU sees, widened to double.
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

static int set_threads(int nthreads)
{
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
    return nthreads;
#else
    (void)nthreads;
    return 1;
#endif
}

int64_t oracle_packed_index(int64_t n, int64_t i, int64_t j)
{
    /* R 'dist' order, pair (i, j) with i > j: column j, row i. */
    if (i < j) { int64_t t = i; i = j; j = t; }
    return j * n - j * (j + 1) / 2 + (i - j - 1);
}

/* One pair, FP64, R semantics.  a and b point at rows of length k. */
double oracle_pair(const float *a, const float *b, int64_t k, int method, double p)
{
    int64_t c;
    if (k <= 0) return NAN;
    switch (method) {
    case ORACLE_EUCLIDEAN: {
        double s = 0.0;
        for (c = 0; c < k; ++c) {
            double t = (double)a[c] - (double)b[c];
            s += t * t;
        }
        return sqrt(s);
    }
    case ORACLE_MAXIMUM: {
        double m = -DBL_MAX;
        for (c = 0; c < k; ++c) {
            double t = fabs((double)a[c] - (double)b[c]);
            if (t > m) m = t;
        }
        return m;
    }
    case ORACLE_MANHATTAN: {
        double s = 0.0;
        for (c = 0; c < k; ++c) s += fabs((double)a[c] - (double)b[c]);
        return s;
    }
    case ORACLE_CANBERRA: {
        double s = 0.0;
        int64_t count = 0;
        for (c = 0; c < k; ++c) {
            double num = fabs((double)a[c] - (double)b[c]);
            double den = fabs((double)a[c] + (double)b[c]);
            if (num > DBL_MIN || den > DBL_MIN) {   /* not 0/0 */
                s += num / den;                      /* x/0 = Inf for a = -b */
                ++count;
            }
        }
        if (count == 0) return NAN;
        if (count != k) s /= ((double)count / (double)k);
        return s;
    }
    case ORACLE_BINARY: {
        int64_t on_any = 0, on_one = 0;
        for (c = 0; c < k; ++c) {
            int x = a[c] != 0.0f, y = b[c] != 0.0f;
            if (x || y) {
                ++on_any;
                if (!(x && y)) ++on_one;
            }
        }
        if (on_any == 0) return 0.0;
        return (double)on_one / (double)on_any;
    }
    case ORACLE_MINKOWSKI: {
        double s = 0.0;
        for (c = 0; c < k; ++c) s += pow(fabs((double)a[c] - (double)b[c]), p);
        return pow(s, 1.0 / p);
    }
    default:
        return NAN;
    }
}

int oracle_dist_packed(const float *x, int64_t n, int64_t k, int64_t ldx,
                       int method, double p, double *out, int nthreads)
{
    int64_t j;
    if (n < 0 || k < 0 || ldx < k || method < 0 || method > 5) return 1;
    if (method == ORACLE_MINKOWSKI && !(p > 0.0 && isfinite(p))) return 1;
    nthreads = set_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 4) num_threads(nthreads)
    for (j = 0; j < n; ++j) {
        int64_t i;
        int64_t base = j * n - j * (j + 1) / 2 - j - 1;
        for (i = j + 1; i < n; ++i)
            out[base + i] = oracle_pair(x + i * ldx, x + j * ldx, k, method, p);
    }
    return 0;
}

int oracle_dist_pairs(const float *x, int64_t k, int64_t ldx,
                      const int64_t *ii, const int64_t *jj, int64_t npairs,
                      int method, double p, double *out, int nthreads)
{
    int64_t q;
    if (k < 0 || ldx < k || method < 0 || method > 5) return 1;
    nthreads = set_threads(nthreads);
#pragma omp parallel for schedule(static) num_threads(nthreads)
    for (q = 0; q < npairs; ++q)
        out[q] = oracle_pair(x + ii[q] * ldx, x + jj[q] * ldx, k, method, p);
    return 0;
}

int oracle_dist_cross(const float *a, int64_t na, const float *b, int64_t nb,
                      int64_t k, int method, double p, double *out, int nthreads)
{
    int64_t y;
    if (na < 0 || nb < 0 || k < 0 || method < 0 || method > 5) return 1;
    nthreads = set_threads(nthreads);
    /* Cross layout of the gputools non-_same kernels (distance.cu:3-38):
       pair (x in a, y in b) at out[y * na + x]. */
#pragma omp parallel for schedule(static) num_threads(nthreads)
    for (y = 0; y < nb; ++y) {
        int64_t xi;
        for (xi = 0; xi < na; ++xi)
            out[y * na + xi] = oracle_pair(a + xi * k, b + y * k, k, method, p);
    }
    return 0;
}

/*
 * Streaming comparator: max relative error of an FP32 packed result against
 * the FP64 oracle, without storing the FP64 matrix.  Pairs whose oracle value
 * is below abs_floor are compared absolutely.  Returns the max error and the
 * (i, j) where it occurs.
 */
double oracle_compare_packed(const float *x, int64_t n, int64_t k, int64_t ldx,
                             int method, double p, const float *got,
                             double abs_floor, int64_t *worst_i, int64_t *worst_j,
                             int nthreads)
{
    double best = 0.0;
    int64_t bi = -1, bj = -1;
    int64_t j;
    nthreads = set_threads(nthreads);
#pragma omp parallel num_threads(nthreads)
    {
        double lbest = 0.0;
        int64_t li = -1, lj = -1;
#pragma omp for schedule(dynamic, 4)
        for (j = 0; j < n; ++j) {
            int64_t i;
            int64_t base = j * n - j * (j + 1) / 2 - j - 1;
            for (i = j + 1; i < n; ++i) {
                double ref = oracle_pair(x + i * ldx, x + j * ldx, k, method, p);
                double g = (double)got[base + i];
                double err;
                if (isnan(ref) || isnan(g)) {
                    err = (isnan(ref) && isnan(g)) ? 0.0 : INFINITY;
                } else if (isinf(ref) || isinf(g)) {
                    err = (ref == g) ? 0.0 : INFINITY;
                } else {
                    double den = fabs(ref) > abs_floor ? fabs(ref) : 1.0;
                    err = fabs(g - ref) / den;
                }
                if (err > lbest) { lbest = err; li = i; lj = j; }
            }
        }
#pragma omp critical
        {
            if (lbest > best) { best = lbest; bi = li; bj = lj; }
        }
    }
    if (worst_i) *worst_i = bi;
    if (worst_j) *worst_j = bj;
    return best;
}

/*
 * Backward of the packed distance vector, FP64:  dX_i = sum_{j != i} G_ij *
 * d(d_ij)/d(x_i), with G the packed dL/dD.
 *   euclidean:  d(d_ij)/dx_i = (x_i - x_j) / d_ij            (0 when d_ij = 0)
 *   minkowski:  d(d_ij)/dx_ic = sign(t_c) |t_c|^(p-1) / d_ij^(p-1), t = x_i - x_j
 * The forward formulas are distance.cu:40-78 and :453-493; gputools has no
 * backward, so this is the mathematics of those formulas.
 */
int oracle_dist_backward(const float *x, int64_t n, int64_t k, int method, double p,
                         const double *g, double *dx, int nthreads)
{
    int64_t i;
    if (method != ORACLE_EUCLIDEAN && method != ORACLE_MINKOWSKI) return 1;
    if (method == ORACLE_MINKOWSKI && !(p >= 1.0 && isfinite(p))) return 1;
    nthreads = set_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 4) num_threads(nthreads)
    for (i = 0; i < n; ++i) {
        int64_t j, c;
        double *row = dx + i * k;
        for (c = 0; c < k; ++c) row[c] = 0.0;
        for (j = 0; j < n; ++j) {
            double d, gij;
            if (j == i) continue;
            gij = g[oracle_packed_index(n, i, j)];
            if (gij == 0.0) continue;
            d = oracle_pair(x + i * k, x + j * k, k, method, p);
            if (!(d > 0.0)) continue;
            if (method == ORACLE_EUCLIDEAN) {
                for (c = 0; c < k; ++c)
                    row[c] += gij * ((double)x[i * k + c] - (double)x[j * k + c]) / d;
            } else {
                double dp = pow(d, p - 1.0);
                for (c = 0; c < k; ++c) {
                    double t = (double)x[i * k + c] - (double)x[j * k + c];
                    double s = (t > 0) - (t < 0);
                    row[c] += gij * s * pow(fabs(t), p - 1.0) / dp;
                }
            }
        }
    }
    return 0;
}

/*
 * Metric-MDS stress S = sum_{i<j} (||y_i - y_j|| - delta_ij)^2 and its
 * gradient dS/dY (FP64), delta packed in 'dist' order.
 */
double oracle_mds_stress(const float *y, int64_t n, int64_t q, const double *delta,
                         double *dy, int nthreads)
{
    double total = 0.0;
    int64_t i;
    nthreads = set_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 4) reduction(+:total) num_threads(nthreads)
    for (i = 0; i < n; ++i) {
        int64_t j, c;
        if (dy) for (c = 0; c < q; ++c) dy[i * q + c] = 0.0;
        for (j = 0; j < n; ++j) {
            double d, r;
            if (j == i) continue;
            d = oracle_pair(y + i * q, y + j * q, q, ORACLE_EUCLIDEAN, 2.0);
            r = d - delta[oracle_packed_index(n, i, j)];
            if (j > i) total += r * r;
            if (dy && d > 0.0)
                for (c = 0; c < q; ++c)
                    dy[i * q + c] += 2.0 * r * ((double)y[i * q + c] - (double)y[j * q + c]) / d;
        }
    }
    return total;
}
