// gt_backward.hip -- gradients of the all-pairs distance vector.
//
// gputools has no backward; these kernels differentiate the forward formulas
// of inst/cuda/distance.cu (euclidean :40-78, minkowski :453-493).  For the
// packed distance vector D and an upstream gradient G = dL/dD (packed, R
// 'dist' order):
//   dL/dx_ic = sum_{j != i} G_ij * dd_ij/dx_ic,
//   euclidean   dd_ij/dx_ic = (x_ic - x_jc) / d_ij
//   minkowski   dd_ij/dx_ic = sign(t) |t|^(p-1) / d_ij^(p-1),  t = x_ic - x_jc
// and a pair with d_ij == 0 contributes 0 (subgradient).
//
// Row-owner sweep: a workgroup owns 64 rows i and one block of up to 128
// features, and visits a share of the 64-row blocks j.  Per (i, j) tile it
// recomputes d_ij exactly (FP32 differences, two-level summation over LDS
// chunks of up to 32 features; D is never stored), forms
// w_ij = G_ij / d_ij^(p-1) in LDS, then accumulates dX_i over its features.
// The 64 x 64 block of packed G is loaded coalesced (lanes along the
// contiguous index of R's 'dist' order, which is i below the diagonal and j
// above it) and staged through LDS.
// When there are few row blocks, the j blocks are split over JS workgroups
// per row block, which write partial dX into a workspace; a second kernel
// adds the JS partials in a fixed order.  No atomics: deterministic.
//
// The same kernel computes the metric-MDS stress S = sum_{i<j} (d_ij - delta_ij)^2
// of an embedding Y and its gradient (G_ij = 2 (d_ij - delta_ij)); S is summed
// in FP64 from per-workgroup partials (the only FP64 on the device).
#include "gt_common.h"
#include "gt_internal.h"

namespace gt {
namespace {

constexpr int BI = 64;          // owned rows per workgroup
constexpr int BJ = 64;          // rows per visited block
constexpr int BTH = 256;        // threads
constexpr int DKC = 32;         // features per distance chunk
constexpr int DLD = BI + 4;     // LDS row length of the k-major distance chunks
constexpr int KBMAX = 128;      // largest feature block per workgroup
constexpr int WLD = BI + 4;     // LDS row length of the G / weight tile (float4 rows)

enum { MODE_GRAD = 0, MODE_MDS = 1 };
enum { BW_EUCLID = 0, BW_MINK3 = 1, BW_MINKP = 2 };

struct BwArgs {
    const float *x;
    int64_t n, k, ldx;
    const float *g;        // MODE_GRAD: packed dL/dD; MODE_MDS: packed delta
    float *dx;
    int64_t lddx;
    float *dxp;            // JS > 1: JS partial dX slices, n x k each
    double *partials;      // MODE_MDS: one per (row block, j split) workgroup
    float p, pm1;          // p and p - 1
};

template <int OP>
__device__ __forceinline__ float dterm(float t)
{
    // contribution of one feature difference to the distance accumulator
    if (OP == BW_EUCLID) return t * t;
    const float a = fabsf(t);
    if (OP == BW_MINK3) return a * a * a;
    return 0.0f;  // BW_MINKP handled by the caller (needs p)
}

// KB: features per workgroup (4, 32 or 128, chosen from k on the host).
// blockIdx.z = js of JS = gridDim.z: this workgroup visits j blocks
// js, js + JS, ...; with JS > 1 it writes partial dX to args.dxp.
template <int OP, int MODE, int KB>
__global__ __launch_bounds__(BTH) void backward_kernel(BwArgs args)
{
    constexpr int XLD = KB + 4;   // LDS row length of the x_J feature block (float4 rows)
    constexpr int DK = KB < DKC ? KB : DKC;   // features per distance chunk
    __shared__ float xi_s[DK][DLD];
    __shared__ float xj_s[DK][DLD];
    __shared__ float w_s[BI][WLD];            // G, then the weights
    __shared__ float xb_s[BJ][XLD];
    __shared__ double red_s[BTH / 64];

    const int tid = threadIdx.x;
    const int64_t n = args.n, k = args.k, ldx = args.ldx;
    const int64_t i0 = (int64_t)blockIdx.x * BI;
    const int64_t f0 = (int64_t)blockIdx.y * KB;
    const int64_t fcount = (k - f0) < KB ? (k - f0) : KB;

    // distance phase mapping: 4 x 4 pairs per thread
    const int ty = tid >> 4, tx = tid & 15;
    // accumulation phase mapping (a register-blocked W x_J product): rows
    // rg * RPT + r, features fg * FPT + m of the feature block
    constexpr int RPT = KB >= 32 ? 4 : 1;            // rows per thread
    constexpr int FPT = KB * BI / (BTH * RPT);       // features per thread: 8, 2, 1
    constexpr int FGROUPS = KB / FPT;
    const int rg = tid / FGROUPS, fg = tid % FGROUPS;
    float xi_r[RPT][FPT], acc[RPT][FPT], rowsum[RPT];
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
        rowsum[r] = 0.0f;
        const int64_t gi = i0 + rg * RPT + r;
#pragma unroll
        for (int m = 0; m < FPT; ++m) {
            const int c = fg * FPT + m;
            // euclidean needs x_i only at the end (dX_i = x_i sum_j w_ij - sum_j w_ij x_j)
            xi_r[r][m] = (OP != BW_EUCLID && gi < n && c < fcount) ? args.x[gi * ldx + f0 + c] : 0.0f;
            acc[r][m] = 0.0f;
        }
    }
    double stress = 0.0;

    // the 64 x 64 block of G is element e = tid + 256 q; below the diagonal
    // (i > j) the packed index runs along i, above it along j
    const int js = blockIdx.z, JS = gridDim.z;
    // G of tile j0 into registers: loaded one tile ahead, so its latency
    // overlaps the previous tile's work
    float gq[(BI * BJ) / BTH];
    auto gfetch = [&](int64_t jt) {
        const bool up = jt > i0;
#pragma unroll
        for (int q = 0; q < (BI * BJ) / BTH; ++q) {
            const int e = tid + BTH * q;
            const int li = up ? (e >> 6) : (e & 63), lj = up ? (e & 63) : (e >> 6);
            const int64_t i = i0 + li, j = jt + lj;
            float g = 0.0f;
            if (i < n && j < n && i != j) {
                const int64_t hi = i > j ? i : j, lo = i > j ? j : i;
                g = args.g[packed_base(n, lo) + hi];
            }
            gq[q] = g;
        }
    };
    if ((int64_t)js * BJ < n) gfetch((int64_t)js * BJ);
    for (int64_t j0 = (int64_t)js * BJ; j0 < n; j0 += (int64_t)JS * BJ) {
        const bool upper = j0 > i0;   // every i < j in this tile
        // ---- d_ij for the 64 x 64 tile, exact FP32 differences ------------
        float dsum[4][4];
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) dsum[a][b] = 0.0f;
        for (int64_t kb = 0; kb < k; kb += DK) {
            __syncthreads();
#pragma unroll
            for (int q = 0; q < (BI * DK) / BTH; ++q) {
                const int e = tid + BTH * q;
                const int row = e / DK, kc = e % DK;
                const int64_t kk = kb + kc;
                const int64_t ri = i0 + row, rj = j0 + row;
                // with one chunk (k <= DK) the owned rows' chunk is the same
                // for every tile: staged on the first only
                if (k > DK || j0 == (int64_t)js * BJ)
                    xi_s[kc][row] = (ri < n && kk < k) ? args.x[ri * ldx + kk] : 0.0f;
                xj_s[kc][row] = (rj < n && kk < k) ? args.x[rj * ldx + kk] : 0.0f;
            }
            __syncthreads();
            float part[4][4];
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int b = 0; b < 4; ++b) part[a][b] = 0.0f;
#pragma unroll 4
            for (int kc = 0; kc < DK; ++kc) {
                const float4 vi = *reinterpret_cast<const float4 *>(&xi_s[kc][ty * 4]);
                const float4 vj = *reinterpret_cast<const float4 *>(&xj_s[kc][tx * 4]);
                const float ai[4] = {vi.x, vi.y, vi.z, vi.w};
                const float aj[4] = {vj.x, vj.y, vj.z, vj.w};
#pragma unroll
                for (int a = 0; a < 4; ++a)
#pragma unroll
                    for (int b = 0; b < 4; ++b) {
                        const float t = ai[a] - aj[b];
                        if (OP == BW_MINKP)
                            part[a][b] += exp2f(args.p * log2f(fabsf(t)));
                        else
                            part[a][b] += dterm<OP>(t);
                    }
            }
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int b = 0; b < 4; ++b) dsum[a][b] += part[a][b];
        }
        // ---- G to LDS (the previous tile's weights were read before the
        // distance phase's barriers), then w_ij = G_ij / d_ij^(p-1) ---------
#pragma unroll
        for (int q = 0; q < (BI * BJ) / BTH; ++q) {
            const int e = tid + BTH * q;
            const int li = upper ? (e >> 6) : (e & 63), lj = upper ? (e & 63) : (e >> 6);
            w_s[li][lj] = gq[q];
        }
        if (j0 + (int64_t)JS * BJ < n) gfetch(j0 + (int64_t)JS * BJ);
        __syncthreads();
        float w_r[4][4];
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                const int li = ty * 4 + a, lj = tx * 4 + b;
                const int64_t i = i0 + li, j = j0 + lj;
                float w = 0.0f;
                if (i < n && j < n && i != j) {
                    const float s = dsum[a][b];
                    const float d = (OP == BW_EUCLID) ? sqrtf(s)
                                    : (OP == BW_MINK3) ? cbrtf(s)
                                                       : powf(s, 1.0f / args.p);
                    const float gv = w_s[li][lj];
                    if (d > 0.0f) {
                        float gij;
                        if (MODE == MODE_MDS) {
                            const float r = d - gv;     // gv = delta_ij
                            stress += (double)r * (double)r;
                            gij = 2.0f * r;
                        } else {
                            gij = gv;
                        }
                        // d^(p-1): p = 2 -> d, p = 3 -> d^2
                        const float dp = (OP == BW_EUCLID) ? d
                                         : (OP == BW_MINK3) ? d * d
                                                            : powf(d, args.pm1);
                        w = gij / dp;
                    } else if (MODE == MODE_MDS) {
                        const float r = -gv;            // d == 0
                        stress += (double)r * (double)r;
                    }
                }
                w_r[a][b] = w;
            }
        // every G value is read: store the weights transposed, w_s[j][i]
        __syncthreads();
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) w_s[tx * 4 + b][ty * 4 + a] = w_r[a][b];
        // ---- x_J feature block ----------------------------------------------
        for (int e = tid; e < BJ * KB; e += BTH) {
            const int row = e / KB, c = e % KB;
            const int64_t rj = j0 + row;
            xb_s[row][c] = (rj < n && c < fcount) ? args.x[rj * ldx + f0 + c] : 0.0f;
        }
        __syncthreads();
        // ---- dX_i += sum_j w_ij phi(x_i - x_j) ------------------------------
        const int jmax = (n - j0) < BJ ? (int)(n - j0) : BJ;
        for (int lj = 0; lj < jmax; ++lj) {
            float w[RPT], xj[FPT];
            if (RPT == 4) {
                const float4 v = *reinterpret_cast<const float4 *>(&w_s[lj][rg * 4]);
                w[0] = v.x;
                w[1] = v.y;
                w[2 % RPT] = v.z;
                w[3 % RPT] = v.w;
            } else {
                w[0] = w_s[lj][rg];
            }
            if (FPT % 4 == 0) {
#pragma unroll
                for (int m = 0; m < FPT; m += 4) {
                    const float4 v = *reinterpret_cast<const float4 *>(&xb_s[lj][fg * FPT + m]);
                    xj[m] = v.x;
                    xj[(m + 1) % FPT] = v.y;
                    xj[(m + 2) % FPT] = v.z;
                    xj[(m + 3) % FPT] = v.w;
                }
            } else if (FPT == 2) {
                const float2 v = *reinterpret_cast<const float2 *>(&xb_s[lj][fg * 2]);
                xj[0] = v.x;
                xj[1 % FPT] = v.y;
            } else {
                xj[0] = xb_s[lj][fg];
            }
#pragma unroll
            for (int r = 0; r < RPT; ++r) {
                if (OP == BW_EUCLID) {
                    rowsum[r] += w[r];
#pragma unroll
                    for (int m = 0; m < FPT; ++m) acc[r][m] = fmaf(-w[r], xj[m], acc[r][m]);
                } else {
#pragma unroll
                    for (int m = 0; m < FPT; ++m) {
                        const float t = xi_r[r][m] - xj[m];
                        const float a = fabsf(t);
                        // sign(t) |t|^(p-1); 0 at t == 0 (p >= 1)
                        const float ph = (OP == BW_MINK3) ? a * t
                                         : (a > 0.0f ? copysignf(exp2f(args.pm1 * log2f(a)), t) : 0.0f);
                        acc[r][m] = fmaf(w[r], ph, acc[r][m]);
                    }
                }
            }
        }
    }

    if (args.dx) {
#pragma unroll
        for (int r = 0; r < RPT; ++r) {
            const int64_t gi = i0 + rg * RPT + r;
            if (gi >= n) continue;
            // JS > 1: partial sums to slice js of the workspace (row stride k)
            float *const out = JS > 1 ? args.dxp + (int64_t)js * n * k + gi * k
                                      : args.dx + gi * args.lddx;
#pragma unroll
            for (int m = 0; m < FPT; ++m) {
                const int c = fg * FPT + m;
                if (c < fcount) {
                    const float v = (OP == BW_EUCLID)
                                        ? fmaf(args.x[gi * ldx + f0 + c], rowsum[r], acc[r][m])
                                        : acc[r][m];
                    out[f0 + c] = v;
                }
            }
        }
    }

    if (MODE == MODE_MDS && blockIdx.y == 0) {
        // each unordered pair was seen by both of its row owners: S = sum / 2
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) stress += __shfl_xor(stress, d, 64);
        __syncthreads();
        if ((tid & 63) == 0) red_s[tid >> 6] = stress;
        __syncthreads();
        if (tid == 0) {
            double s = 0.0;
            for (int w = 0; w < BTH / 64; ++w) s += red_s[w];
            args.partials[(int64_t)js * gridDim.x + blockIdx.x] = 0.5 * s;
        }
    }
}

// Sum of the per-workgroup FP64 partials, fixed order (deterministic).
__global__ __launch_bounds__(256) void reduce_partials_kernel(const double *partials, int64_t count,
                                                              double *out)
{
    __shared__ double s_[256];
    double s = 0.0;
    for (int64_t e = threadIdx.x; e < count; e += 256) s += partials[e];
    s_[threadIdx.x] = s;
    __syncthreads();
    for (int d = 128; d >= 1; d >>= 1) {
        if ((int)threadIdx.x < d) s_[threadIdx.x] += s_[threadIdx.x + d];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = s_[0];
}

// dx = sum over the JS partial slices, in slice order.
__global__ __launch_bounds__(256) void sum_partial_dx_kernel(const float *__restrict__ dxp, int64_t n,
                                                             int64_t k, int js_count,
                                                             float *__restrict__ dx, int64_t lddx)
{
    const int64_t total = n * k;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        float s = dxp[e];
        for (int q = 1; q < js_count; ++q) s += dxp[(int64_t)q * total + e];
        dx[(e / k) * lddx + e % k] = s;
    }
}

int feature_block(int64_t k) { return k <= 4 ? 4 : (k <= 32 ? 32 : KBMAX); }

// j splits: about 2048 workgroups in all, at most 16 splits and no more than
// the number of j blocks.
int j_splits(int64_t n, int64_t k)
{
    const int64_t gx = (n + BI - 1) / BI, gy = (k + feature_block(k) - 1) / feature_block(k);
    int64_t js = (2048 + gx * gy - 1) / (gx * gy);
    if (js > 16) js = 16;
    if (js > gx) js = gx;
    return (int)(js < 1 ? 1 : js);
}

template <int MODE, int KB>
hipError_t launch_kb(int op, BwArgs &args, int js, hipStream_t stream)
{
    dim3 grid((unsigned)((args.n + BI - 1) / BI), (unsigned)((args.k + KB - 1) / KB), (unsigned)js);
    if (op == BW_EUCLID)
        hipLaunchKernelGGL((backward_kernel<BW_EUCLID, MODE, KB>), grid, dim3(BTH), 0, stream, args);
    else if (op == BW_MINK3)
        hipLaunchKernelGGL((backward_kernel<BW_MINK3, MODE, KB>), grid, dim3(BTH), 0, stream, args);
    else
        hipLaunchKernelGGL((backward_kernel<BW_MINKP, MODE, KB>), grid, dim3(BTH), 0, stream, args);
    return hipGetLastError();
}

template <int MODE>
hipError_t launch_mode(int op, BwArgs &args, hipStream_t stream)
{
    const int js = j_splits(args.n, args.k);
    hipError_t e;
    if (args.k <= 4)
        e = launch_kb<MODE, 4>(op, args, js, stream);
    else if (args.k <= 32)
        e = launch_kb<MODE, 32>(op, args, js, stream);
    else
        e = launch_kb<MODE, KBMAX>(op, args, js, stream);
    if (e != hipSuccess || js == 1 || args.dx == nullptr) return e;
    int64_t blocks = (args.n * args.k + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(sum_partial_dx_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, args.dxp,
                       args.n, args.k, js, args.dx, args.lddx);
    return hipGetLastError();
}

}  // namespace

int64_t mds_partials_count(int64_t n, int64_t q) { return (n + BI - 1) / BI * j_splits(n, q); }

size_t backward_workspace_bytes(int64_t n, int64_t k)
{
    if (n < 2 || k < 1) return 0;
    const int js = j_splits(n, k);
    return js > 1 ? (size_t)js * (size_t)n * (size_t)k * sizeof(float) : 0;
}

hipError_t launch_backward(int op, const float *x, int64_t n, int64_t k, int64_t ldx, float p,
                           const float *g_packed, float *dx, int64_t lddx, float *workspace,
                           hipStream_t stream)
{
    if (n <= 0 || k <= 0) return hipSuccess;
    BwArgs args;
    args.x = x;
    args.n = n;
    args.k = k;
    args.ldx = ldx;
    args.g = g_packed;
    args.dx = dx;
    args.lddx = lddx;
    args.dxp = workspace;
    args.partials = nullptr;
    args.p = p;
    args.pm1 = p - 1.0f;
    const int bop = (op == OP_EUCLID) ? BW_EUCLID : (op == OP_MINK3) ? BW_MINK3 : BW_MINKP;
    return launch_mode<MODE_GRAD>(bop, args, stream);
}

hipError_t launch_mds(const float *y, int64_t n, int64_t q, int64_t ldy, const float *delta,
                      float *dy, int64_t lddy, double *stress, double *partials, float *workspace,
                      hipStream_t stream)
{
    BwArgs args;
    args.x = y;
    args.n = n;
    args.k = q;
    args.ldx = ldy;
    args.g = delta;
    args.dx = dy;
    args.lddx = lddy;
    args.dxp = workspace;
    args.partials = partials;
    args.p = 2.0f;
    args.pm1 = 1.0f;
    hipError_t e = launch_mode<MODE_MDS>(BW_EUCLID, args, stream);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(reduce_partials_kernel, dim3(1), dim3(256), 0, stream, partials,
                       mds_partials_count(n, q), stress);
    return hipGetLastError();
}

}  // namespace gt
