// gt_hclust.hip -- the clustering stage of gpuDistClust / gpuHclust.
//
// Agglomerative clustering on a full, symmetric n x n FP32 distance matrix
// in HBM (the layout gt_dist_device_full produces), with the ten
// Lance-Williams updates of the reference (inst/cuda/hcluster.cu:106-400) and
// its merge rule (src/hcluster.cpp:40-201): each round merges the closest
// pair (sub < sup) over the lower triangle, the surviving cluster keeps index
// sup, and its row and column are rewritten with the method's formula.
//
// The reference rescans the whole lower triangle every round (find_min1,
// n/1024 launches of n^2/2 reads in total per round) and pays three kernel
// launches plus a module load per launch and round.  Here:
//   - hclust_init_kernel (many workgroups, one wave per row) puts +inf on the
//     diagonal, sets every count to 1 and computes each row's minimum over
//     columns 0..r-1 (value and column);
//   - hclust_merge_kernel (one persistent workgroup of 1024 threads) runs all
//     n - 1 rounds.  Per round: the argmin over the cached row minima, the
//     update of row and column sup (row and column sub become +inf), then the
//     cached minima are repaired: row sup is rescanned; a row r > sup takes
//     the new D[r][sup] if it beats its cached minimum, and is rescanned only
//     when that minimum sat at column sub or sup and the new value does not
//     settle it; a row sub < r < sup is rescanned only when its minimum sat
//     at column sub.  Rows whose minimum is unaffected cost one comparison.
//   - up to n = 9216 the per-row state (counts, cached minima and the new row
//     sup) stays in LDS for the whole run.  One CU issues every store, and
//     its scattered column stores (one cache line each) dominate a round, so
//     this variant writes only row and column sup: a merged-away column is
//     recognised by its negative count instead of being set to +inf.  Only
//     the end of a round waits for the stores; until then the rescans take
//     row sup and column sup from the LDS copy.  A rescanned row is split
//     over several waves when few rows need one.
// Ties are broken exactly as the reference's 32-lane reductions do
// (hcluster.cu:27-57, 67-90): smallest value, then the lane (index mod 32)
// that wins the reduction tree (bit-reversed lane order), then smallest
// index; +inf and NaN entries are never candidates.  The update
// formulas keep the reference's arithmetic types (FP32, with the FP64
// promotions its double constants cause for median, flexible and flexible
// group) and round every product separately (no FMA contraction), so merges and heights
// match the reference's loop bit for bit.  When no finite distance is left
// (undefined in the reference) the two lowest-index clusters merge at +inf.
#include "gt_common.h"
#include "gt_internal.h"

namespace gt {
namespace {

constexpr int HTHREADS = 1024;
constexpr int HWAVES = HTHREADS / 64;
constexpr int RESCAN_CAP = 2048;   // rows listed for rescan per round (LDS)
constexpr int HB = 4;              // rows per thread per batch (repair)
constexpr int HBU = 4;             // columns per thread per batch (update)
constexpr int NO_COL = -1;         // row without a candidate
constexpr int PENDING = -2;        // row queued for a rescan
constexpr int HC_LDS_ROWS = 9216;  // largest n whose row state fits in LDS (16 B per row)

struct HclustArgs {
    float *d;
    int64_t ldd;
    int n;
    int method;
    float lambda, beta;
    float *count;      // > 0: cluster size; < 0: merged away (minus its size)
    float *minv;       // cached row minimum over columns 0..r-1
    int *minc;         // its column, NO_COL or PENDING
    int *sub, *sup;
    float *height;
    bool vec4;         // rows start 16 B aligned
};

// Order of equal values in the reference's 32-lane reductions: lane t keeps
// the first of its indices t, t+32, ...; the tree then folds lanes t and
// t + 16, t + 8, ..., t + 1 into t, keeping t on ties, so lane priority is
// the 5-bit reversal of the lane number (even lanes before odd ones, ...).
__device__ __forceinline__ unsigned tie_key(int idx)
{
    return ((__builtin_bitreverse32((unsigned)idx) >> 27) << 26) | ((unsigned)idx >> 5);
}

// (v, i) beats (bv, bi); bv = +inf with bi = NO_COL is the empty candidate
__device__ __forceinline__ bool beats(float v, int i, float bv, int bi)
{
    return v < bv || (v == bv && bi != NO_COL && tie_key(i) < tie_key(bi));
}

// wave-wide argmin of (value, index); lanes without a candidate hold
// (+inf, NO_COL)
__device__ __forceinline__ void wave_argmin(float &v, int &i)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        const float ov = __shfl_xor(v, off, 64);
        const int oi = __shfl_xor(i, off, 64);
        if (oi != NO_COL && (i == NO_COL || beats(ov, oi, v, i))) {
            v = ov;
            i = oi;
        }
    }
}

__device__ __forceinline__ void take(float v, int c, float &bv, int &bi)
{
    if (v < INFINITY && (bi == NO_COL || beats(v, c, bv, bi))) {
        bv = v;
        bi = c;
    }
}

// argmin of row r over columns 0..r-1 by one wave (all lanes return it).
// Rows that start 16 B aligned are read 16 B per lane, four loads in flight.
__device__ __forceinline__ void scan_row(const float *row, int r, bool vec4, float &bv, int &bi)
{
    const int lane = threadIdx.x & 63;
    bv = INFINITY;
    bi = NO_COL;
    int c = 0;
    if (vec4) {
        const int full = r & ~1023;   // columns covered by whole 4 x 256 blocks
        for (; c < full; c += 1024) {
            float4 q[4];
#pragma unroll
            for (int u = 0; u < 4; ++u)
                q[u] = *reinterpret_cast<const float4 *>(row + c + u * 256 + lane * 4);
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int cc = c + u * 256 + lane * 4;
                take(q[u].x, cc, bv, bi);
                take(q[u].y, cc + 1, bv, bi);
                take(q[u].z, cc + 2, bv, bi);
                take(q[u].w, cc + 3, bv, bi);
            }
        }
    }
    for (int cc = c + lane; cc < r; cc += 64) take(row[cc], cc, bv, bi);
    wave_argmin(bv, bi);
}

// This lane's argmin of row[c] over c in [c0, c1) (c0 a multiple of 4)
// among live columns (count[c] > 0: entries of merged-away clusters are not
// kept at +inf), with column rw taking fixv: the value the current round has
// stored there, which may still be in flight.
__device__ __forceinline__ void scan_cols(const float *row, int c0, int c1, int rw, float fixv,
                                          const float *count, bool vec4, float &bv, int &bi)
{
    const int lane = threadIdx.x & 63;
    auto put = [&](float v, int c) {
        if (c < c1 && count[c] > 0.0f) take(c == rw ? fixv : v, c, bv, bi);
    };
    if (!vec4) {
        for (int cc = c0 + lane; cc < c1; cc += 64) put(row[cc], cc);
        return;
    }
    // 2048 columns per batch: eight 16 B loads per lane in flight
    for (int c = c0; c < c1; c += 2048) {
        float4 q[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int cc = c + u * 256 + lane * 4;
            q[u] = make_float4(INFINITY, INFINITY, INFINITY, INFINITY);
            if (cc + 3 < c1) {
                q[u] = *reinterpret_cast<const float4 *>(row + cc);
            } else if (cc < c1) {
                q[u].x = row[cc];
                if (cc + 1 < c1) q[u].y = row[cc + 1];
                if (cc + 2 < c1) q[u].z = row[cc + 2];
            }
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int cc = c + u * 256 + lane * 4;
            put(q[u].x, cc);
            put(q[u].y, cc + 1);
            put(q[u].z, cc + 2);
            put(q[u].w, cc + 3);
        }
    }
}

// Barrier over LDS traffic only: this wave's global stores may still be in
// flight when the others pass it.
__device__ __forceinline__ void lds_barrier()
{
    asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

// A value the compiler cannot see through: a product passed through it is
// rounded on its own and never fused into the following add (the library is
// built with -ffp-contract=fast, under which contraction pragmas are ignored).
__device__ __forceinline__ float sep(float x)
{
    asm volatile("" : "+v"(x));
    return x;
}
__device__ __forceinline__ double sep(double x)
{
    asm volatile("" : "+v"(x));
    return x;
}

// Lance-Williams update of the reference's kernels (hcluster.cu:106-400):
// bot = D[sup][c], top = D[sub][c]; nr = new size of sup, np = size of sub,
// nq = old size of sup, nk = count[c], val = this round's merge height.
// Every product is rounded separately, as the reference's unfused code does.
template <int METHOD>
__device__ __forceinline__ float lw_update(float bot, float top, float nr, float np, float nq,
                                           float nk, float val, float lambda, float beta)
{
    switch (METHOD) {
    case HC_SINGLE:
        return fminf(bot, top);
    case HC_COMPLETE:
        return fmaxf(bot, top);
    case HC_WPGMA:
    case HC_MCQUITTY:
        return (float)((double)(bot + top) / 2.0);
    case HC_AVERAGE:
        return (sep(top * np) + sep(bot * nq)) / nr;
    case HC_MEDIAN:
        return (float)((double)(bot + top) / 2.0 - (double)val / 4.0);
    case HC_CENTROID:
        return (sep(top * np) + sep(bot * nq)) / nr - (np * nq * val) / (nr * nr);
    case HC_FLEXIBLE_GROUP:
        return (float)((sep((double)bot * (1.0 - (double)lambda) * (double)np) +
                        sep((double)top * (1.0 - (double)lambda) * (double)nq)) /
                           (double)nr +
                       (double)sep(beta * val));
    case HC_FLEXIBLE:
        return (float)((sep((double)bot * (1.0 - (double)lambda)) +
                        sep((double)top * (1.0 - (double)lambda))) /
                           2.0 +
                       (double)sep(beta * val));
    case HC_WARD: {
        if (nr == -nk) return INFINITY;
        float b = (sep(bot * (np + nk)) + sep(top * (nq + nk)) - sep(val * nk));
        b /= (nr + nk);
        return isinf(b) ? INFINITY : b;
    }
    default:
        return fminf(bot, top);
    }
}

// Diagonal to +inf, counts to 1, and every row's minimum: one wave per row.
__global__ __launch_bounds__(256) void hclust_init_kernel(HclustArgs a)
{
    const int lane = threadIdx.x & 63;
    const int64_t waves = (int64_t)gridDim.x * 4;
    for (int64_t r = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6); r < a.n; r += waves) {
        float *row = a.d + r * a.ldd;
        if (lane == 0) {
            row[r] = INFINITY;
            a.count[r] = 1.0f;
        }
        float bv;
        int bi;
        scan_row(row, (int)r, a.vec4, bv, bi);
        if (lane == 0) {
            a.minv[r] = bv;
            a.minc[r] = bi;
        }
    }
}

// LDSRES: the per-row state (counts, cached minima and the rewritten row
// sup) lives in LDS for the whole run, n <= HC_LDS_ROWS; otherwise in the
// global workspace.
template <int METHOD, bool LDSRES>
__global__ __launch_bounds__(HTHREADS) void hclust_merge_kernel(HclustArgs a)
{
    extern __shared__ float hs_dyn[];
    __shared__ float red_v[HWAVES];
    __shared__ int red_i[HWAVES];
    __shared__ int s_cw, s_rw, s_nres, s_overflow;
    __shared__ float s_val;
    __shared__ int rescan[RESCAN_CAP];
    __shared__ float part_v[HWAVES];
    __shared__ int part_i[HWAVES];

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int n = a.n;
    const int64_t ldd = a.ldd;
    float *cntp, *minvp, *newp;
    int *mincp;
    if constexpr (LDSRES) {
        cntp = hs_dyn;
        minvp = hs_dyn + n;
        mincp = reinterpret_cast<int *>(hs_dyn + 2 * n);
        newp = hs_dyn + 3 * n;   // D[sup][c] as rewritten this round
        for (int r = tid; r < n; r += HTHREADS) {
            cntp[r] = a.count[r];
            minvp[r] = a.minv[r];
            mincp[r] = a.minc[r];
        }
        __syncthreads();
    } else {
        cntp = a.count;
        minvp = a.minv;
        mincp = a.minc;
        newp = nullptr;
    }
    for (int iter = 0; iter + 1 < n; ++iter) {
        // ---- the closest pair: argmin over the cached row minima ----------
        float bv = INFINITY;
        int bi = NO_COL;
        for (int r = tid; r < n; r += HTHREADS) {
            const float v = minvp[r];
            if (mincp[r] >= 0 && v < INFINITY && (bi == NO_COL || beats(v, r, bv, bi))) {
                bv = v;
                bi = r;
            }
        }
        wave_argmin(bv, bi);
        if (lane == 0) {
            red_v[wave] = bv;
            red_i[wave] = bi;
        }
        if constexpr (LDSRES)
            lds_barrier();
        else
            __syncthreads();
        // the waves' candidates reduced by wave 0 (the order on (value, tie
        // key) is total, so the lane order of the reduction does not matter)
        float v = INFINITY;
        int r = NO_COL;
        if (wave == 0) {
            if (lane < HWAVES) {
                v = red_v[lane];
                r = red_i[lane];
            }
            wave_argmin(v, r);
        }
        if (tid == 0) {
            int cw, rw;
            if (r != NO_COL) {
                rw = r;
                cw = mincp[r];
            } else {
                // no finite distance left: the two lowest-index clusters
                cw = -1;
                rw = -1;
                for (int i = 0; i < n && rw < 0; ++i)
                    if (cntp[i] > 0.0f) {
                        if (cw < 0)
                            cw = i;
                        else
                            rw = i;
                    }
                v = INFINITY;
            }
            a.height[iter] = v;
            a.sub[iter] = cw;
            a.sup[iter] = rw;
            cntp[rw] += cntp[cw];
            cntp[cw] *= -1.0f;
            s_cw = cw;
            s_rw = rw;
            s_val = v;
            s_nres = 0;
            s_overflow = 0;
        }
        if constexpr (LDSRES)
            lds_barrier();
        else
            __syncthreads();
        const int cw = s_cw, rw = s_rw;
        const float val = s_val;
        const float nr = cntp[rw];
        const float np = -cntp[cw];
        const float nq = nr - np;

        // ---- rewrite row/column sup; row/column sub become +inf ----------
        float *const dcw = a.d + (int64_t)cw * ldd;
        float *const drw = a.d + (int64_t)rw * ldd;
        // batches of HBU columns per thread: every load of a batch is issued
        // before its stores, which could otherwise alias them
        // (8 columns per thread in the LDS variant: one batch of loads at
        // n = 8192, none queued behind the first batch's column stores)
        constexpr int HBC = LDSRES ? 2 * HBU : HBU;
        for (int c0 = tid; c0 < n; c0 += HTHREADS * HBC) {
            float bot[HBC], top[HBC], nk[HBC];
#pragma unroll
            for (int u = 0; u < HBC; ++u) {
                const int c = c0 + u * HTHREADS;
                if (c < n && c != cw && c != rw) {
                    bot[u] = drw[c];
                    top[u] = dcw[c];
                    if constexpr (!LDSRES) nk[u] = cntp[c];
                }
            }
#pragma unroll
            for (int u = 0; u < HBC; ++u) {
                const int c = c0 + u * HTHREADS;
                if (c >= n) break;
                float v = INFINITY;
                if constexpr (LDSRES) {
                    // row and column sub are left as they are (sub is dead,
                    // and scans skip dead columns); a dead column c holds a
                    // stale value, so its new entry is +inf.  A dead row or
                    // column is never read again: only live c are stored.
                    const float nkc = cntp[c];
                    if (c != cw && c != rw && nkc > 0.0f)
                        v = lw_update<METHOD>(bot[u], top[u], nr, np, nq, nkc, val, a.lambda, a.beta);
                    newp[c] = v;
                    if (nkc > 0.0f) {
                        drw[c] = v;
                        a.d[(int64_t)c * ldd + rw] = v;
                    }
                } else {
                    if (c != cw && c != rw)
                        v = lw_update<METHOD>(bot[u], top[u], nr, np, nq, nk[u], val, a.lambda, a.beta);
                    dcw[c] = INFINITY;
                    a.d[(int64_t)c * ldd + cw] = INFINITY;
                    drw[c] = v;
                    a.d[(int64_t)c * ldd + rw] = v;
                }
            }
        }
        if constexpr (LDSRES)
            lds_barrier();
        else
            __syncthreads();

        // ---- repair the cached row minima --------------------------------
        for (int r0 = cw + tid; r0 < n; r0 += HTHREADS * HB) {
            float cnt[HB], mvs[HB], fs[HB];
            int mcs[HB];
#pragma unroll
            for (int u = 0; u < HB; ++u) {
                const int r = r0 + u * HTHREADS;
                if (r < n) {
                    cnt[u] = cntp[r];
                    mvs[u] = minvp[r];
                    mcs[u] = mincp[r];
                    fs[u] = LDSRES ? newp[r] : drw[r];
                }
            }
#pragma unroll
            for (int u = 0; u < HB; ++u) {
            const int r = r0 + u * HTHREADS;
            if (r >= n) break;
            bool redo = false;
            if (r == cw) {
                minvp[r] = INFINITY;
                mincp[r] = NO_COL;
                continue;
            }
            if (cnt[u] <= 0.0f) continue;
            const int mc = mcs[u];
            const float mv = mvs[u];
            if (r == rw) {
                redo = true;
            } else if (r < rw) {
                redo = (mc == cw);
            } else {
                const float f = fs[u];   // the new D[r][sup]
                if (mc == cw) {
                    redo = !(f < mv);
                    if (!redo) {
                        minvp[r] = f;
                        mincp[r] = rw;
                    }
                } else if (mc == rw) {
                    redo = !(f <= mv);
                    if (!redo) minvp[r] = f;
                } else if (f < INFINITY && (mc == NO_COL || beats(f, rw, mv, mc))) {
                    minvp[r] = f;
                    mincp[r] = rw;
                }
            }
            if (redo) {
                mincp[r] = PENDING;
                const int slot = atomicAdd(&s_nres, 1);
                if (slot < RESCAN_CAP)
                    rescan[slot] = r;
                else
                    s_overflow = 1;
            }
            }
        }
        if constexpr (LDSRES)
            lds_barrier();
        else
            __syncthreads();

        // ---- rescans ------------------------------------------------------
        const int nres = s_nres < RESCAN_CAP ? s_nres : RESCAN_CAP;
        if constexpr (LDSRES) {
            // This round's stores to row and column sup may still be in
            // flight: row sup is read from its LDS copy, and in other rows
            // column sup takes the LDS copy's value.  Dead columns (sub's
            // included) are skipped by their count.
            const float *const dr = a.d;
            if (nres < HWAVES) {
                // few rows: each is split over HWAVES / nres waves, whose
                // partial argmins are combined (the order of (value, tie key)
                // is total, so any grouping gives the same minimum)
                const int S = nres > 0 ? HWAVES / nres : 1;
                const int q = wave / S, sg = wave % S;
                float v = INFINITY;
                int c = NO_COL;
                if (q < nres) {
                    const int r = rescan[q];
                    const int per = (((r + 2047) >> 11) + S - 1) / S;   // 2048-column batches
                    const int c0 = min(r, sg * per * 2048), c1 = min(r, (sg + 1) * per * 2048);
                    if (r == rw) {
                        for (int cc = c0 + lane; cc < c1; cc += 64) take(newp[cc], cc, v, c);
                    } else {
                        scan_cols(dr + (int64_t)r * ldd, c0, c1, rw, newp[r], cntp, a.vec4, v, c);
                    }
                }
                wave_argmin(v, c);
                if (lane == 0) {
                    part_v[wave] = v;
                    part_i[wave] = c;
                }
                lds_barrier();
                if (tid < nres) {
                    float bv = INFINITY;
                    int bi = NO_COL;
                    for (int w = tid * S; w < tid * S + S; ++w)
                        if (part_i[w] != NO_COL && (bi == NO_COL || beats(part_v[w], part_i[w], bv, bi))) {
                            bv = part_v[w];
                            bi = part_i[w];
                        }
                    minvp[rescan[tid]] = bv;
                    mincp[rescan[tid]] = bi;
                }
            } else {
                for (int q = wave; q < nres; q += HWAVES) {
                    const int r = rescan[q];
                    float v = INFINITY;
                    int c = NO_COL;
                    if (r == rw) {
                        for (int cc = lane; cc < r; cc += 64) take(newp[cc], cc, v, c);
                    } else {
                        scan_cols(dr + (int64_t)r * ldd, 0, r, rw, newp[r], cntp, a.vec4, v, c);
                    }
                    wave_argmin(v, c);
                    if (lane == 0) {
                        minvp[r] = v;
                        mincp[r] = c;
                    }
                }
                if (s_overflow) {
                    // more rows than the list holds: find the rest by their
                    // marker (every listed row has been rescanned by now)
                    lds_barrier();
                    for (int r = cw + 1 + wave; r < n; r += HWAVES) {
                        if (mincp[r] != PENDING) continue;
                        float v = INFINITY;
                        int c = NO_COL;
                        if (r == rw) {
                            for (int cc = lane; cc < r; cc += 64) take(newp[cc], cc, v, c);
                        } else {
                            scan_cols(dr + (int64_t)r * ldd, 0, r, rw, newp[r], cntp, a.vec4, v, c);
                        }
                        wave_argmin(v, c);
                        if (lane == 0) {
                            minvp[r] = v;
                            mincp[r] = c;
                        }
                    }
                }
            }
            // this round's global stores complete before the next round
            // reads rows sup and sub
            __syncthreads();
            continue;
        }
        for (int q = wave; q < nres; q += HWAVES) {
            const int r = rescan[q];
            float v;
            int c;
            scan_row(a.d + (int64_t)r * ldd, r, a.vec4, v, c);
            if (lane == 0) {
                minvp[r] = v;
                mincp[r] = c;
            }
        }
        if (s_overflow) {
            // more rows than the list holds: find the rest by their marker
            for (int r = cw + 1 + wave; r < n; r += HWAVES) {
                if (mincp[r] != PENDING) continue;
                float v;
                int c;
                scan_row(a.d + (int64_t)r * ldd, r, a.vec4, v, c);
                if (lane == 0) {
                    minvp[r] = v;
                    mincp[r] = c;
                }
            }
        }
        __syncthreads();
    }
}

}  // namespace

hipError_t launch_hclust(float *d, int64_t n, int64_t ldd, int method, float lambda, float beta,
                         float *count, float *minv, int *minc, int *sub, int *sup, float *height,
                         int state, hipStream_t stream)
{
    if (n < 2) return hipSuccess;
    HclustArgs a;
    a.d = d;
    a.ldd = ldd;
    a.n = (int)n;
    a.method = method;
    a.lambda = lambda;
    a.beta = beta;
    a.count = count;
    a.minv = minv;
    a.minc = minc;
    a.sub = sub;
    a.sup = sup;
    a.height = height;
    a.vec4 = (ldd % 4 == 0) && ((reinterpret_cast<uintptr_t>(d) & 15) == 0);
    int64_t blocks = (n + 3) / 4;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(hclust_init_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    // 16 B of LDS per row (count, cached minimum and its column, the new row
    // sup) next to the 8 KB of static LDS: 155 KB at the largest n
    static_assert(HC_LDS_ROWS == HC_LDS_MAX_N, "LDS row limit");
    if (state == HC_STATE_LDS && n > HC_LDS_ROWS) return hipErrorInvalidValue;
    const bool lds = state == HC_STATE_LDS || (state == HC_STATE_AUTO && n <= HC_LDS_ROWS);
    const unsigned shm = lds ? (unsigned)(16 * n) : 0u;
    switch (method) {
#define GT_HC_CASE(m)                                                                          \
    case m:                                                                                    \
        if (lds) {                                                                             \
            e = hipFuncSetAttribute(reinterpret_cast<const void *>(hclust_merge_kernel<m, true>), \
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);     \
            if (e != hipSuccess) return e;                                                     \
            hipLaunchKernelGGL((hclust_merge_kernel<m, true>), dim3(1), dim3(HTHREADS), shm,   \
                               stream, a);                                                     \
        } else {                                                                               \
            hipLaunchKernelGGL((hclust_merge_kernel<m, false>), dim3(1), dim3(HTHREADS), 0,    \
                               stream, a);                                                     \
        }                                                                                      \
        break;
    GT_HC_CASE(HC_SINGLE)
    GT_HC_CASE(HC_COMPLETE)
    GT_HC_CASE(HC_WPGMA)
    GT_HC_CASE(HC_AVERAGE)
    GT_HC_CASE(HC_MEDIAN)
    GT_HC_CASE(HC_CENTROID)
    GT_HC_CASE(HC_FLEXIBLE_GROUP)
    GT_HC_CASE(HC_FLEXIBLE)
    GT_HC_CASE(HC_WARD)
    GT_HC_CASE(HC_MCQUITTY)
#undef GT_HC_CASE
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace gt
