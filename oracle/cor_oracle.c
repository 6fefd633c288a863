/*
 * oracle/cor_oracle.c -- CPU restatement of gputools' Pearson correlation
 * (gpuCor(method = "pearson"), its t-scores and gpuTtest).  TEST
 * INFRASTRUCTURE ONLY: loaded by tests/ and bench.py's CPU leg, never by the
 * product library.
 *
 * Layout, as gpuCor passes it (R/gpuCor.R:33-37, src/rinterface.cpp:33-55):
 * variable x of A is the dim consecutive floats at a + x * dim (a column of
 * R's sample x variable matrix); results are na x nb row-major, pair (x, y)
 * at [x * nb + y] (correlation.cu:48-50, 145-149).
 *
 * oracle_pmcc follows the three kernels of pmcc (src/correlation.cpp:66-160)
 * in FP64, per pair (x, y):
 *   use 0 "everything"           (gpuMeansNoTest / gpuSDNoTest / gpuPMCCNoTest,
 *                                 inst/cuda/correlation.cu:152-290): every
 *                                 sample is used, n = dim;
 *   use 1 "pairwise.complete.obs" (gpuMeans / gpuSD / gpuPMCC,
 *                                 correlation.cu:7-150): samples where both
 *                                 values are non-NaN, n = their count.
 *   mean_a = sum a / n                              (correlation.cu:46-49)
 *   sd_a   = sqrt(sum (a - mean_a)^2 / (n - 1))     (correlation.cu:97-100)
 *   corr   = sum ((a - mean_a)/sd_a)((b - mean_b)/sd_b) / (n - 1)
 *                                                   (correlation.cu:133-149)
 * The literal formulas give the reference's special values: n = 0 gives
 * 0 / (0 - 1) = -0.0, n = 1 or a constant variable gives NaN.
 *
 * oracle_signif restates the t-scores in FP32, operation by operation (this
 * file is compiled with -ffp-contract=off):
 *   cap 1: hostSignif (src/correlation.cpp:51-64), what pmcc returns:
 *          corr >= 0.999 -> 10000, else corr * sqrt((n - 2) / (1 - corr^2));
 *   cap 0: gpuSignif (inst/cuda/correlation.cu:292-310), what gpuTtest
 *          returns: the same formula without the cap.
 * (The reference's GPU kernels are compiled with fast math; the restatement
 * uses IEEE division and square root.)
 *
 * oracle_kendall follows gpuKendall (inst/cuda/kendall.cu:4-47), what
 * gpuCor(method = "kendall") returns (RgpuKendall, src/rinterface.cpp:128-132):
 * a sample pair i < j is concordant when x_j > x_i and y_j > y_i, or
 * x_j < x_i and y_j < y_i, or x_j == x_i and y_j == y_i (kendall.cu:18-22;
 * comparisons with NaN are false); then, in FP32 (kendall.cu:40-45),
 *   denom = n (n - 1) / 2, discordant = denom - concordant,
 *   numer = concordant - discordant, tau = (double) numer / (double) denom.
 * The count is exact (64-bit); the reference sums it in FP32, which is the
 * same below 2^24 concordant pairs.  Output nb x na, pair (x, y) at
 * [y * na + x] (kendall.cu:45).
 */
#include <math.h>
#include <stdint.h>

#include "oracle.h"

static double pair_corr(const float *a, const float *b, int64_t dim, int use, double *count)
{
    double n = 0.0, sa = 0.0, sb = 0.0;
    for (int64_t k = 0; k < dim; ++k) {
        if (use && (isnan(a[k]) || isnan(b[k]))) continue;
        sa += a[k];
        sb += b[k];
        n += 1.0;
    }
    const double ma = sa / n, mb = sb / n;
    double qa = 0.0, qb = 0.0;
    for (int64_t k = 0; k < dim; ++k) {
        if (use && (isnan(a[k]) || isnan(b[k]))) continue;
        qa += (a[k] - ma) * (a[k] - ma);
        qb += (b[k] - mb) * (b[k] - mb);
    }
    const double sda = sqrt(qa / (n - 1.0)), sdb = sqrt(qb / (n - 1.0));
    double s = 0.0;
    for (int64_t k = 0; k < dim; ++k) {
        if (use && (isnan(a[k]) || isnan(b[k]))) continue;
        s += ((a[k] - ma) / sda) * ((b[k] - mb) / sdb);
    }
    *count = n;
    return s / (n - 1.0);
}

int oracle_pmcc(int use, const float *a, int64_t na, const float *b, int64_t nb, int64_t dim,
                double *pairs, double *corr, int nthreads)
{
    if (use != 0 && use != 1) return 1;
    if (nthreads < 1) nthreads = 1;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
    for (int64_t x = 0; x < na; ++x)
        for (int64_t y = 0; y < nb; ++y) {
            double n;
            corr[x * nb + y] = pair_corr(a + x * dim, b + y * dim, dim, use, &n);
            pairs[x * nb + y] = n;
        }
    return 0;
}

void oracle_signif(const float *pairs, const float *corr, int64_t n, int cap, float *ts)
{
    for (int64_t i = 0; i < n; ++i) {
        const float cor = corr[i];
        if (cap && cor >= 0.999f) {
            ts[i] = 10000.0f;
        } else {
            const float c2 = cor * cor;
            const float den = 1.0f - c2;
            const float rad = (pairs[i] - 2.0f) / den;
            ts[i] = cor * sqrtf(rad);
        }
    }
}

static uint64_t concordant_pairs(const float *a, const float *b, int64_t n)
{
    uint64_t c = 0;
    for (int64_t i = 0; i < n; ++i)
        for (int64_t j = i + 1; j < n; ++j)
            c += (uint64_t)(((a[j] > a[i]) && (b[j] > b[i])) + ((a[j] < a[i]) && (b[j] < b[i])) +
                            ((a[j] == a[i]) && (b[j] == b[i])));
    return c;
}

int oracle_kendall(const float *a, int64_t na, const float *b, int64_t nb, int64_t n, double *tau,
                   uint64_t *count, int nthreads)
{
    if (nthreads < 1) nthreads = 1;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
    for (int64_t y = 0; y < nb; ++y)
        for (int64_t x = 0; x < na; ++x) {
            const uint64_t c = concordant_pairs(a + x * n, b + y * n, n);
            const float concordant = (float)c;
            float denom = (float)n;
            denom = (denom * (denom - 1.f)) / 2.f;
            const float discordant = denom - concordant;
            const float numer = concordant - discordant;
            tau[y * na + x] = (double)numer / (double)denom;
            if (count) count[y * na + x] = c;
        }
    return 0;
}
