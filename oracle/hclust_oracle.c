/*
 * oracle/hclust_oracle.c -- CPU restatement of gputools' hierarchical
 * clustering stage (gpuDistClust / gpuHclust).  TEST INFRASTRUCTURE ONLY:
 * loaded by tests/ and bench.py's checks, never by the product library.
 *
 * It follows the reference literally, including its arithmetic types and the
 * tie order of its 32-thread reductions, so that merges and heights can be
 * compared bit for bit:
 *   - hclusterPreparedDistances (src/hcluster.cpp:40-201): every point starts
 *     as a cluster of count 1; n - 1 rounds of (row minima, global minimum,
 *     Lance-Williams update of the surviving row and column);
 *   - convert_kernel (inst/cuda/hcluster.cu:4-9): the diagonal becomes +inf;
 *   - find_min1_kernel (hcluster.cu:11-58): for every row r with count > 0,
 *     the minimum over columns 0..r.  Thread t of 32 scans columns t, t+32,
 *     ... keeping the first strictly smaller value; a tree reduction then
 *     keeps the left operand unless it is strictly greater.  Rows with
 *     count < 0 get (+inf, 0);
 *   - find_min2_kernel (hcluster.cu:60-104): the same reduction over the row
 *     minima; the winning row r is sup, its column is sub, the value is the
 *     merge height; count[sup] += count[sub]; count[sub] *= -1;
 *   - the update kernels (hcluster.cu:106-400): for every column c,
 *     top = D[sub][c], bot = D[sup][c]; bot becomes the method's formula
 *     (+inf for c = sub or sup), top becomes +inf, and both are written to the
 *     row and the column of sub and sup.  The formulas keep the reference's C
 *     types: float arithmetic, except where a double constant promotes the
 *     expression (median, flexible, flexible group);
 *   - formatClustering, getPrintOrder, depthFirst (src/rinterface.cpp:
 *     241-292): R's hclust merge matrix (column-major, n-1 x 2) and order.
 *
 * Method ids follow hc_method (src/hcluster.h:5-16).  The reference leaves
 * the winner undefined when every remaining distance is +inf or NaN (its
 * column index is read from uninitialised shared memory); here that case
 * merges the two lowest-index active clusters at height +inf.
 *
 * Pinned against scipy.cluster.hierarchy.linkage (tests/golden) for single,
 * complete, average, weighted (= wpgma, mcquitty), median and centroid.
 * flexible, flexible group and ward follow the reference's own formulas
 * (ward weights the surviving row by the absorbed cluster's size and vice
 * versa), which no library computes: parity unpinned beyond invariants.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>

#include "oracle.h"

#define HC_THREADS 32

/* The reference's 32-lane argmin: per-lane strict-less scan, then a tree
 * reduction that replaces the left value only when it is strictly greater. */
static void hc_argmin(const float *v, int64_t stride_v, int64_t count, float *best, int64_t *where)
{
    float vals[HC_THREADS];
    int64_t idx[HC_THREADS];
    for (int t = 0; t < HC_THREADS; ++t) {
        vals[t] = INFINITY;
        idx[t] = -1;
        for (int64_t i = t; i < count; i += HC_THREADS) {
            const float x = v[i * stride_v];
            if (x < vals[t]) {
                vals[t] = x;
                idx[t] = i;
            }
        }
    }
    for (int s = HC_THREADS / 2; s > 0; s >>= 1)
        for (int t = 0; t < s; ++t)
            if (vals[t] > vals[t + s]) {
                vals[t] = vals[t + s];
                idx[t] = idx[t + s];
            }
    *best = vals[0];
    *where = idx[0];
}

static float hc_update(int method, float bot, float top, float nr, float np, float nq, float nk,
                       float val, float lambda, float beta)
{
    switch (method) {
    case ORACLE_HC_SINGLE:
        return fminf(bot, top);
    case ORACLE_HC_COMPLETE:
        return fmaxf(bot, top);
    case ORACLE_HC_WPGMA:
    case ORACLE_HC_MCQUITTY:
        return (float)((double)(bot + top) / 2.0);
    case ORACLE_HC_AVERAGE:
        return (top * np + bot * nq) / nr;
    case ORACLE_HC_MEDIAN:
        return (float)((double)(bot + top) / 2.0 - (double)val / 4.0);
    case ORACLE_HC_CENTROID:
        return (top * np + bot * nq) / nr - (np * nq * val) / (nr * nr);
    case ORACLE_HC_FLEXIBLE_GROUP:
        return (float)(((double)bot * (1.0 - (double)lambda) * (double)np +
                        (double)top * (1.0 - (double)lambda) * (double)nq) / (double)nr +
                       (double)(beta * val));
    case ORACLE_HC_FLEXIBLE:
        return (float)(((double)bot * (1.0 - (double)lambda) + (double)top * (1.0 - (double)lambda)) /
                           2.0 +
                       (double)(beta * val));
    case ORACLE_HC_WARD: {
        if (nr == -nk) return INFINITY;
        float b = (bot * (np + nk) + top * (nq + nk) - val * nk);
        b /= (nr + nk);
        if (isinf(b)) b = INFINITY;
        return b;
    }
    default:
        return fminf(bot, top);
    }
}

int oracle_hclust(float *d, int64_t n, int64_t ldd, int method, float lambda, float beta,
                  int *sub, int *sup, float *height)
{
    if (n < 1 || ldd < n || method < 0 || method > ORACLE_HC_MCQUITTY) return -1;
    float *count = (float *)malloc((size_t)n * sizeof(float));
    float *minv = (float *)malloc((size_t)n * sizeof(float));
    int64_t *minc = (int64_t *)malloc((size_t)n * sizeof(int64_t));
    if (!count || !minv || !minc) {
        free(count);
        free(minv);
        free(minc);
        return -2;
    }
    for (int64_t i = 0; i < n; ++i) {
        count[i] = 1.0f;
        d[i * ldd + i] = INFINITY;
    }
    for (int64_t iter = 0; iter + 1 < n; ++iter) {
        for (int64_t r = 0; r < n; ++r) {
            if (count[r] < 0.0f) {
                minv[r] = INFINITY;
                minc[r] = 0;
            }
            if (count[r] <= 0.0f) continue;
            hc_argmin(d + r * ldd, 1, r + 1, &minv[r], &minc[r]);
        }
        float v;
        int64_t rw;
        hc_argmin(minv, 1, n, &v, &rw);
        int64_t cw;
        if (rw >= 0 && minc[rw] >= 0) {
            cw = minc[rw];
        } else {
            /* no finite distance left: the two lowest-index active clusters */
            cw = -1;
            rw = -1;
            for (int64_t i = 0; i < n && rw < 0; ++i)
                if (count[i] > 0.0f) {
                    if (cw < 0)
                        cw = i;
                    else
                        rw = i;
                }
            v = INFINITY;
        }
        height[iter] = v;
        sub[iter] = (int)cw;
        sup[iter] = (int)rw;
        count[rw] += count[cw];
        count[cw] *= -1.0f;
        const float nr = count[rw];
        const float np = (float)(-1.0 * (double)count[cw]);
        const float nq = nr - np;
        for (int64_t c = 0; c < n; ++c) {
            float top = d[cw * ldd + c];
            float bot = d[rw * ldd + c];
            bot = hc_update(method, bot, top, nr, np, nq, count[c], v, lambda, beta);
            if (c == cw || c == rw) bot = INFINITY;
            top = INFINITY;
            d[cw * ldd + c] = top;
            d[c * ldd + cw] = top;
            d[rw * ldd + c] = bot;
            d[c * ldd + rw] = bot;
        }
    }
    free(count);
    free(minv);
    free(minc);
    return 0;
}

static void hc_depth_first(int len, const int *merge, int level, int *otop, int *order)
{
    const int left = level, right = level + len;
    if (merge[right] < 0) {
        order[*otop] = -merge[right];
        (*otop)--;
    } else {
        hc_depth_first(len, merge, merge[right] - 1, otop, order);
    }
    if (merge[left] < 0) {
        order[*otop] = -merge[left];
        (*otop)--;
    } else {
        hc_depth_first(len, merge, merge[left] - 1, otop, order);
    }
}

int oracle_hclust_format(int len, const int *sub, const int *sup, int *merge, int *order)
{
    if (len < 0) return -1;
    for (int i = 0; i < len; ++i) {
        merge[i] = -(sub[i] + 1);
        merge[i + len] = -(sup[i] + 1);
    }
    for (int i = 0; i < len; ++i) {
        for (int j = i + 1; j < len; ++j) {
            if ((merge[j] == merge[i]) || (merge[j] == merge[i + len])) merge[j] = i + 1;
            if ((merge[j + len] == merge[i]) || (merge[j + len] == merge[i + len])) merge[j + len] = i + 1;
            if (((merge[j + len] < 0) && (merge[j] > 0)) ||
                ((merge[j] > 0) && (merge[j + len] > 0) && (merge[j] > merge[j + len]))) {
                const int h = merge[j];
                merge[j] = merge[j + len];
                merge[j + len] = h;
            }
        }
    }
    if (len > 0) {
        int otop = len;
        hc_depth_first(len, merge, len - 1, &otop, order);
    } else {
        order[0] = 1;
    }
    return 0;
}
