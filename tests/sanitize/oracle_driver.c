/* Runs every oracle function on small inputs (edge cases included) in a
 * build with AddressSanitizer and UndefinedBehaviorSanitizer; any invalid
 * access, leak or undefined operation aborts the program.  Results are
 * checked loosely (symmetry, zero distances, known values): the numerical
 * checks live in the Python tests. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../oracle/oracle.h"

#define CHECK(c)                                                             \
    do {                                                                     \
        if (!(c)) {                                                          \
            fprintf(stderr, "check failed at line %d: %s\n", __LINE__, #c);  \
            return 1;                                                        \
        }                                                                    \
    } while (0)

static float frand(unsigned *s)
{
    *s = *s * 1103515245u + 12345u;
    return (float)((*s >> 8) & 0xffff) / 65536.0f - 0.5f;
}

static int distances(void)
{
    unsigned seed = 7;
    const int64_t sizes[][2] = {{0, 3}, {1, 3}, {2, 1}, {5, 0}, {17, 33}, {64, 9}};
    for (size_t c = 0; c < sizeof(sizes) / sizeof(sizes[0]); ++c) {
        const int64_t n = sizes[c][0], k = sizes[c][1];
        float *x = malloc((size_t)(n * k + 1) * sizeof(float));
        for (int64_t i = 0; i < n * k; ++i) x[i] = frand(&seed);
        if (n > 3 && k > 1) {
            x[2 * k] = 0.0f;            /* zeros and an exact duplicate row */
            memcpy(x + 3 * k, x + 1 * k, (size_t)k * sizeof(float));
        }
        const int64_t np = n > 1 ? n * (n - 1) / 2 : 0;
        double *out = malloc((size_t)(np + 1) * sizeof(double));
        for (int m = ORACLE_EUCLIDEAN; m <= ORACLE_MINKOWSKI; ++m) {
            CHECK(oracle_dist_packed(x, n, k, k, m, 3.0, out, 2) == 0);
            for (int64_t e = 0; e < np; ++e) CHECK(isnan(out[e]) || out[e] >= 0.0);
            if (n > 3 && k > 1 && m != ORACLE_BINARY)
                CHECK(out[oracle_packed_index(n, 3, 1)] == 0.0);
        }
        if (n >= 2 && k > 0) {
            int64_t ii[3] = {1, n - 1, 0}, jj[3] = {0, 0, n - 1};
            double pv[3];
            CHECK(oracle_dist_pairs(x, k, k, ii, jj, 3, ORACLE_EUCLIDEAN, 2.0, pv, 1) == 0);
            CHECK(pv[1] == pv[2]);
            double *cross = malloc((size_t)(n * n) * sizeof(double));
            CHECK(oracle_dist_cross(x, n, x, n, k, ORACLE_MANHATTAN, 1.0, cross, 2) == 0);
            for (int64_t i = 0; i < n; ++i) CHECK(cross[i * n + i] == 0.0);
            free(cross);
            float *got = malloc((size_t)(np + 1) * sizeof(float));
            CHECK(oracle_dist_packed(x, n, k, k, ORACLE_EUCLIDEAN, 2.0, out, 1) == 0);
            for (int64_t e = 0; e < np; ++e) got[e] = (float)out[e];
            int64_t wi = -1, wj = -1;
            const double err = oracle_compare_packed(x, n, k, k, ORACLE_EUCLIDEAN, 2.0, got,
                                                     1e-30, &wi, &wj, 2);
            CHECK(err >= 0.0 && err < 1e-6);
            free(got);
            double *g = malloc((size_t)(np + 1) * sizeof(double));
            double *dx = malloc((size_t)(n * k) * sizeof(double));
            for (int64_t e = 0; e < np; ++e) g[e] = 1.0;
            CHECK(oracle_dist_backward(x, n, k, ORACLE_MINKOWSKI, 3.0, g, dx, 2) == 0);
            const double s = oracle_mds_stress(x, n, k, out, dx, 2);
            CHECK(fabs(s) < 1e-12);     /* delta equals the embedding's distances */
            free(g);
            free(dx);
        }
        free(out);
        free(x);
    }
    return 0;
}

static int clustering(void)
{
    unsigned seed = 11;
    for (int n = 1; n <= 40; n += 13) {
        float *pts = malloc((size_t)n * 3 * sizeof(float));
        for (int i = 0; i < n * 3; ++i) pts[i] = frand(&seed);
        double *dp = malloc((size_t)(n * n + 1) * sizeof(double));
        CHECK(oracle_dist_packed(pts, n, 3, 3, ORACLE_EUCLIDEAN, 2.0, dp, 1) == 0);
        for (int m = ORACLE_HC_SINGLE; m <= ORACLE_HC_MCQUITTY; ++m) {
            float *d = malloc((size_t)n * n * sizeof(float));
            for (int i = 0; i < n; ++i)
                for (int j = 0; j < n; ++j)
                    d[i * n + j] = i == j ? 0.0f
                                          : (float)dp[oracle_packed_index(n, i > j ? i : j, i > j ? j : i)];
            int *sub = malloc((size_t)(n + 1) * sizeof(int)), *sup = malloc((size_t)(n + 1) * sizeof(int));
            float *h = malloc((size_t)(n + 1) * sizeof(float));
            CHECK(oracle_hclust(d, n, n, m, 0.5f, 0.5f, sub, sup, h) == 0);
            int *merge = malloc((size_t)(2 * n + 2) * sizeof(int)), *order = malloc((size_t)(n + 1) * sizeof(int));
            CHECK(oracle_hclust_format(n - 1, sub, sup, merge, order) == 0);
            for (int i = 0; i + 1 < n; ++i) CHECK(sub[i] < sup[i]);
            free(merge);
            free(order);
            free(sub);
            free(sup);
            free(h);
            free(d);
        }
        free(dp);
        free(pts);
    }
    return 0;
}

static int correlations(void)
{
    unsigned seed = 13;
    const int na = 9, nb = 5, dim = 37;
    float *a = malloc(na * dim * sizeof(float)), *b = malloc(nb * dim * sizeof(float));
    for (int i = 0; i < na * dim; ++i) a[i] = frand(&seed);
    for (int i = 0; i < nb * dim; ++i) b[i] = frand(&seed);
    a[3] = NAN;
    b[dim + 4] = NAN;
    double *pairs = malloc(na * nb * sizeof(double)), *corr = malloc(na * nb * sizeof(double));
    for (int use = 0; use <= 1; ++use) CHECK(oracle_pmcc(use, a, na, b, nb, dim, pairs, corr, 2) == 0);
    float *fp = malloc(na * nb * sizeof(float)), *fc = malloc(na * nb * sizeof(float)),
          *ts = malloc(na * nb * sizeof(float));
    for (int i = 0; i < na * nb; ++i) {
        fp[i] = (float)pairs[i];
        fc[i] = (float)corr[i];
    }
    oracle_signif(fp, fc, na * nb, 1, ts);
    oracle_signif(fp, fc, na * nb, 0, ts);
    double *tau = malloc(na * na * sizeof(double));   /* both calls: at most na x na */
    uint64_t *count = malloc(na * na * sizeof(uint64_t));
    CHECK(oracle_kendall(a, na, b, nb, dim, tau, count, 2) == 0);
    CHECK(oracle_kendall(a, na, a, na, 1, tau, count, 1) == 0);
    free(a);
    free(b);
    free(pairs);
    free(corr);
    free(fp);
    free(fc);
    free(ts);
    free(tau);
    free(count);
    return 0;
}

int main(void)
{
    if (distances() || clustering() || correlations()) return 1;
    printf("oracle sanitizer run: ok\n");
    return 0;
}
