// gt_pearson.hip -- Pearson correlation of gpuCor(method = "pearson") on gfx950.
//
// Replaces pmcc (src/correlation.cpp:66-160) and its kernels
// (inst/cuda/correlation.cu:7-290).  gputools launches three kernels over an
// na x nb grid of 16-thread blocks, one block per pair, each re-reading both
// variables from global memory (means, then standard deviations, then the sum
// of standard-score products), and computes the t-scores on the host
// (correlation.cpp:51-64).
//
// use = "everything": every pair uses every sample, so means and standard
// deviations belong to the variables, not to the pairs.  A pre-pass
// standardises each variable once,
//     z = (v - mean) / sqrt(sum (v - mean)^2),
// and the correlation of a pair is the dot product z_x . z_y, which equals
// the reference's sum of standard-score products / (n - 1).  The dot products
// of a 256 x 256 tile of pairs run on the matrix cores with split-precision
// operands (gt_mfma.h), or in FP32 on the VALU (the plain distance tile with a
// dot-product term).
//
// use = "pairwise.complete.obs": a pair uses the samples where both values
// are present, so means and deviations are per pair.  A VALU tile kernel keeps
// the reference's two passes per pair: counts and sums give the pair's means,
// then the centred second moments give the correlation
//     sum d_x d_y / (sqrt(sum d_x^2) sqrt(sum d_y^2)).
//
// Every path writes pairs, corr and hostSignif's t-score in one epilogue.
// With a == b (gpuCor(x)) only the tiles on and below the diagonal are
// computed and each value is stored at (x, y) and (y, x), so the result is
// exactly symmetric, as the reference's is.
#include "gt_common.h"
#include "gt_internal.h"
#include "gt_mfma.h"

#include <math.h>

namespace gt {
namespace {

// t-score of a correlation over npairs samples, operation by operation as
// hostSignif (src/correlation.cpp:51-64, CAP) or gpuSignif
// (inst/cuda/correlation.cu:292-310, no cap) compute it in FP32: IEEE
// division and square root, cor^2 rounded before the subtraction.
template <bool CAP>
__device__ __forceinline__ float t_score(float npairs, float cor)
{
    if (CAP && cor >= 0.999f) return 10000.0f;
    float c2 = cor * cor;
    asm volatile("" : "+v"(c2));
    return cor * sqrtf((npairs - 2.0f) / (1.0f - c2));
}

__device__ __forceinline__ float wave_sum(float v)
{
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
    return v;
}

__device__ __forceinline__ float wave_max(float v)
{
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v = fmaxf(v, __shfl_xor(v, d, 64));
    return v;
}

// ---------------------------------------------------------------------------
// Pre-pass of use = "everything": one wave per variable.
// ---------------------------------------------------------------------------
enum { PREP_BF16 = SPLIT_BF16, PREP_FP16 = SPLIT_FP16, PREP_FP32 = 2 };

// z = (v - mean) / sqrt(sum (v - mean)^2), the mean and the sum over all dim
// samples (NaN in, NaN out; a constant variable gives 0 / 0 = NaN, as the
// reference's zero standard deviation does).  PREP_FP32 writes z (n x dim);
// the split forms write the terms of z, term t of variable r at
// terms[(t * n + r) * kp + c], samples dim..kp-1 zero.  FP16 scales z by a
// power of two that brings max |z| into [2^14, 2^15) and records the inverse.
template <int PREP>
__global__ __launch_bounds__(256) void pmcc_prep_kernel(const float *__restrict__ x, int64_t n,
                                                        int64_t dim, int64_t ldx, int64_t kp,
                                                        unsigned short *__restrict__ terms,
                                                        float *__restrict__ iscale,
                                                        float *__restrict__ z)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int64_t row = (int64_t)blockIdx.x * 4 + wave; row < n; row += (int64_t)gridDim.x * 4) {
        const float *xr = x + row * ldx;
        float s = 0.0f;
        for (int64_t c = lane; c < dim; c += 64) s += xr[c];
        const float mean = wave_sum(s) / (float)dim;
        float q = 0.0f, mx = 0.0f;
        for (int64_t c = lane; c < dim; c += 64) {
            const float d = xr[c] - mean;
            q = fmaf(d, d, q);
            mx = fmaxf(mx, fabsf(d));
        }
        const float root = sqrtf(wave_sum(q));
        mx = wave_max(mx);
        if constexpr (PREP == PREP_FP32) {
            for (int64_t c = lane; c < dim; c += 64) z[row * dim + c] = (xr[c] - mean) / root;
        } else {
            float sc = 1.0f;
            if (PREP == PREP_FP16) {
                const float m = mx / root;   // max |z|
                int e = 0;
                if (m > 0.0f && m < INFINITY) {
                    e = ilogbf(m);
                    e = e < -100 ? -100 : (e > 100 ? 100 : e);
                }
                sc = ldexpf(1.0f, 14 - e);
                if (lane == 0) iscale[row] = ldexpf(1.0f, e - 14);
            }
            constexpr int NT = Split<PREP>::NT;
            for (int64_t c0 = (int64_t)lane * 8; c0 < kp; c0 += 512) {
                unsigned short t[8][NT];
#pragma unroll
                for (int c = 0; c < 8; ++c) {
                    const float v = (c0 + c < dim) ? ((xr[c0 + c] - mean) / root) * sc : 0.0f;
                    Split<PREP>::split(v, t[c]);
                }
#pragma unroll
                for (int q2 = 0; q2 < NT; ++q2) {
                    uint4 w;
                    w.x = t[0][q2] | ((unsigned)t[1][q2] << 16);
                    w.y = t[2][q2] | ((unsigned)t[3][q2] << 16);
                    w.z = t[4][q2] | ((unsigned)t[5][q2] << 16);
                    w.w = t[6][q2] | ((unsigned)t[7][q2] << 16);
                    *reinterpret_cast<uint4 *>(terms + ((int64_t)q2 * n + row) * kp + c0) = w;
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// MFMA tile of use = "everything": corr(x, y) = z_x . z_y.
// ---------------------------------------------------------------------------
constexpr int MBM = 256;        // tile side (variables)
constexpr int MBK = 16;         // samples per main-loop chunk (one MFMA k-step)
constexpr int MKPAD = 32;       // split terms are zero-padded to a multiple of this
constexpr int MROWB = MBK * 2;  // bytes of one row of one term in one chunk
constexpr int MTHREADS = 1024;  // 16 waves, 4 (x) x 4 (y), 64 x 64 pairs each
constexpr int MSUPER = 2;       // super-tile side in tiles (L2 locality)
constexpr int MTERM = MBM * MROWB;   // one term of one side, one chunk (8 KB)
constexpr int MNSTAGE = 3;

struct PmccTileArgs {
    const unsigned short *ta, *tb;   // split terms, NT x na x kp and NT x nb x kp
    int64_t na, nb, kp;
    const float *isa, *isb;          // FP16: inverse power-of-two scales
    float *corr, *ts, *pairs;
    int64_t ldo;
    float fpairs;                    // samples per pair (dim)
    int64_t super_b;                 // super-tiles along y (rectangular schedule)
    int64_t tiles_a, tiles_b;
    int64_t id_offset, ids_this;
};

// SYM: a == b; tiles with J <= I, values mirrored to (y, x).  VEC: ldo and the
// output pointers allow 16-byte mirrored stores.
template <int SPLIT, bool SYM, bool VEC>
__global__ __launch_bounds__(MTHREADS) void pmcc_tile_kernel(PmccTileArgs args)
{
    typedef typename Split<SPLIT>::vec vec;
    constexpr int NT = Split<SPLIT>::NT;
    constexpr int STAGE = 2 * NT * MTERM;
    // after the main loop the stages hold each wave's transpose region for
    // the mirrored stores: 64 rows of MTRL floats
    constexpr int MTRL = 36;
    constexpr int TRBYTES = (MTHREADS / 64) * 64 * MTRL * 4;
    constexpr int SBYTES = MNSTAGE * STAGE > TRBYTES ? MNSTAGE * STAGE : TRBYTES;
    __shared__ __attribute__((aligned(1024))) unsigned char smem[SBYTES + 2 * MBM * 4];
    float *iA = reinterpret_cast<float *>(smem + SBYTES);
    float *iB = iA + MBM;

    const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wr = wave >> 2, wc = wave & 3;   // x rows wr*64.., y rows wc*64..

    // Super-tiled schedule: super-tile st covers MSUPER x MSUPER tiles; tiles
    // past the edge (or above the diagonal when SYM) exit at once.
    const int64_t id = args.id_offset + xcd_remap(blockIdx.x, args.ids_this);
    const int64_t st = id / (MSUPER * MSUPER);
    const int w = (int)(id % (MSUPER * MSUPER));
    int64_t SI, SJ;
    if (SYM) {
        int si, sj;
        tri_tile(st, si, sj);
        SI = si;
        SJ = sj;
    } else {
        SI = st / args.super_b;
        SJ = st % args.super_b;
    }
    const int64_t I = SI * MSUPER + w / MSUPER;
    const int64_t J = SJ * MSUPER + w % MSUPER;
    if (I >= args.tiles_a || J >= args.tiles_b || (SYM && J > I)) return;
    const int64_t x0 = I * MBM, y0 = J * MBM;
    const int64_t na = args.na, nb = args.nb, kp = args.kp;

    if (SPLIT == SPLIT_FP16) {
        if (tid < MBM) {
            const int64_t r = x0 + tid;
            iA[tid] = r < na ? args.isa[r] : 1.0f;
        } else if (tid < 2 * MBM) {
            const int64_t r = y0 + (tid - MBM);
            iB[tid - MBM] = r < nb ? args.isb[r] : 1.0f;
        }
    }

    // ---- main loop: direct global -> LDS loads, MNSTAGE-deep, MFMA products
    // LDS chunk image: for stage g, side s (0 = x, 1 = y) and term t, 256 rows
    // x 32 B; the two 16 B halves of row R are swapped when
    // ((R >> 2) ^ (R >> 3)) & 1, so that the b128 operand reads of 8 or 16
    // consecutive rows hit distinct banks.  global_load_lds writes lane l's
    // 16 B at the wave's LDS base + 16 l: waves 0-7 fill side x and waves 8-15
    // side y, wave w rows 32(w%8)..32(w%8)+31 of every term block of its side.
    // Rows past the end are clamped to the last row: they only feed pairs
    // that are never stored.
    typedef __attribute__((address_space(3))) void lds_void_t;
    typedef const __attribute__((address_space(1))) void gbl_void_t;
    const int lside = wave >> 3, lrb = wave & 7;
    const unsigned short *gsrc;
    int64_t term_stride;
    {
        const int row = lrb * 32 + (lane >> 1);
        const int half = (lane & 1) ^ (((row >> 2) ^ (row >> 3)) & 1);
        const int64_t nside = lside ? nb : na;
        const int64_t r = (lside ? y0 : x0) + row;
        gsrc = (lside ? args.tb : args.ta) + (r < nside ? r : nside - 1) * kp + half * 8;
        term_stride = nside * kp;
    }
    unsigned char *const lbase = smem + lrb * 1024;
    auto issue = [&](int chunk) {
        const int g = chunk % MNSTAGE;
        const int64_t kb = (int64_t)chunk * MBK;
#pragma unroll
        for (int t = 0; t < NT; ++t)
            __builtin_amdgcn_global_load_lds(
                (gbl_void_t *)(gsrc + t * term_stride + kb),
                (lds_void_t *)(lbase + g * STAGE + (lside * NT + t) * MTERM), 16, 0, 0);
    };
    const int lrow = lane & 31;
    const int roff = lrow * MROWB + (((lane >> 5) ^ (((lrow >> 2) ^ (lrow >> 3)) & 1)) * 16);

    f32x16_t acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[a][b][r] = 0.0f;

    const int nchunks = (int)(kp / MBK);
#pragma unroll
    for (int c = 0; c < MNSTAGE - 1; ++c)
        if (c < nchunks) issue(c);
    for (int c = 0; c < nchunks; ++c) {
        // chunk c landed when at most the next chunk's NT loads per wave are
        // outstanding (vector memory loads complete in order)
        if (c + 1 < nchunks) {
            if (NT == 3)
                asm volatile("s_waitcnt vmcnt(3)" ::: "memory");
            else
                asm volatile("s_waitcnt vmcnt(2)" ::: "memory");
        } else {
            asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        }
        // every wave's part of chunk c is in LDS and every wave finished chunk
        // c-1, whose stage the next issue overwrites (a bare s_barrier:
        // __syncthreads() would also drain the later chunks' loads)
        __builtin_amdgcn_s_barrier();
        asm volatile("" ::: "memory");
        if (c + MNSTAGE - 1 < nchunks) issue(c + MNSTAGE - 1);
        const unsigned char *sg = smem + (c % MNSTAGE) * STAGE + roff;
        vec av[NT][2];
#pragma unroll
        for (int t = 0; t < NT; ++t)
#pragma unroll
            for (int a = 0; a < 2; ++a)
                av[t][a] = *reinterpret_cast<const vec *>(sg + t * MTERM + (wr * 64 + a * 32) * MROWB);
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            vec bv[NT];
#pragma unroll
            for (int t = 0; t < NT; ++t)
                bv[t] = *reinterpret_cast<const vec *>(sg + (NT + t) * MTERM +
                                                       (wc * 64 + b * 32) * MROWB);
#pragma unroll
            for (int a = 0; a < 2; ++a) {
                acc[a][b] = Split<SPLIT>::mfma(av[0][a], bv[0], acc[a][b]);
                acc[a][b] = Split<SPLIT>::mfma(av[0][a], bv[1], acc[a][b]);
                acc[a][b] = Split<SPLIT>::mfma(av[1][a], bv[0], acc[a][b]);
                if (NT == 3) {
                    acc[a][b] = Split<SPLIT>::mfma(av[0][a], bv[NT - 1], acc[a][b]);
                    acc[a][b] = Split<SPLIT>::mfma(av[NT - 1][a], bv[0], acc[a][b]);
                    acc[a][b] = Split<SPLIT>::mfma(av[1][a], bv[1], acc[a][b]);
                }
            }
        }
    }
    if (SPLIT == SPLIT_FP16) __syncthreads();   // iA / iB visible (written before the loop)

    // ---- epilogue: corr, t-score, pairs ------------------------------------
    // 32x32 accumulator layout: element r of lane l is x row 8(r/4) + 4(l/32)
    // + r%4 and y column l%32, so lanes 0..31 store 32 consecutive y of one x.
    const int ly0 = wc * 64 + (lane & 31);
    const float fp = args.fpairs;
    const int64_t ldo = args.ldo;
    float sb[2];
#pragma unroll
    for (int b = 0; b < 2; ++b) sb[b] = (SPLIT == SPLIT_FP16) ? iB[ly0 + b * 32] : 1.0f;
    // SYM: off-diagonal tiles store (x, y) and (y, x); diagonal tiles store
    // x >= y and mirror x > y, so the two halves are the same values
    const bool diag = SYM && (I == J);
    // whole off-diagonal tiles mirror through LDS: each wave writes its 64 x
    // 32 block (one output, one a) transposed into its own region, then
    // stores it back as rows of 32 consecutive x (128 B per row and store),
    // instead of 16-byte pieces of 32 rows
    const bool fastm = SYM && VEC && !diag && x0 + MBM <= na && y0 + MBM <= nb;
    float *const tr = reinterpret_cast<float *>(smem) + wave * 64 * MTRL;
    // lane's four x (8 rq + 4 (lane >> 5) ..) of row (lane & 31) + 32 b
    auto tr_put = [&](int rq, int b, const float (&v)[4]) {
        *reinterpret_cast<float4 *>(tr + ((lane & 31) + 32 * b) * MTRL + 8 * rq + 4 * (lane >> 5)) =
            make_float4(v[0], v[1], v[2], v[3]);
    };
    // the wave's transposed 64 x 32 block out as rows of 32 x
    auto tr_store = [&](float *out, int a) {
#pragma unroll
        for (int it = 0; it < 8; ++it) {
            const int ly = 8 * it + (lane >> 3), lx = (lane & 7) * 4;
            const float4 v = *reinterpret_cast<const float4 *>(tr + ly * MTRL + lx);
            *reinterpret_cast<float4 *>(out + (y0 + wc * 64 + ly) * ldo + x0 + wr * 64 + a * 32 + lx) = v;
        }
    };
#pragma unroll
    for (int a = 0; a < 2; ++a) {
        if (fastm) __syncthreads();   // the region's earlier reads (or the main loop's) are done
        float tva[4][2][4];
#pragma unroll
        for (int rq = 0; rq < 4; ++rq) {
            // the four x rows of this lane: lx0 .. lx0 + 3
            const int lx0 = wr * 64 + a * 32 + 8 * rq + 4 * (lane >> 5);
            float cv[2][4];
            float (&tv)[2][4] = tva[rq];
#pragma unroll
            for (int r4 = 0; r4 < 4; ++r4) {
                const int64_t x = x0 + lx0 + r4;
                const float sj = (SPLIT == SPLIT_FP16) ? iA[lx0 + r4] : 1.0f;
#pragma unroll
                for (int b = 0; b < 2; ++b) {
                    float g = acc[a][b][rq * 4 + r4];
                    // undo the scales one at a time (their product can overflow)
                    if (SPLIT == SPLIT_FP16) g = (g * sj) * sb[b];
                    cv[b][r4] = g;
                    tv[b][r4] = t_score<true>(fp, g);
                    const int64_t y = y0 + ly0 + b * 32;
                    if (x < na && y < nb && (!diag || x >= y)) {
                        const int64_t o = x * ldo + y;
                        args.corr[o] = g;
                        args.ts[o] = tv[b][r4];
                        args.pairs[o] = fp;
                    }
                }
            }
            if (fastm) {
#pragma unroll
                for (int b = 0; b < 2; ++b) tr_put(rq, b, cv[b]);
            } else if (SYM) {
                // (y, x) for the four x of this lane: one 16-byte store per
                // output where the rows allow it
#pragma unroll
                for (int b = 0; b < 2; ++b) {
                    const int64_t y = y0 + ly0 + b * 32;
                    const int64_t x = x0 + lx0;
                    if (y >= nb) continue;
                    const int64_t o = y * ldo + x;
                    if (VEC && !diag && x + 3 < na) {
                        *reinterpret_cast<float4 *>(args.corr + o) =
                            make_float4(cv[b][0], cv[b][1], cv[b][2], cv[b][3]);
                        *reinterpret_cast<float4 *>(args.ts + o) =
                            make_float4(tv[b][0], tv[b][1], tv[b][2], tv[b][3]);
                        *reinterpret_cast<float4 *>(args.pairs + o) = make_float4(fp, fp, fp, fp);
                    } else {
#pragma unroll
                        for (int r4 = 0; r4 < 4; ++r4) {
                            if (x + r4 >= na || (diag && x + r4 <= y)) continue;
                            args.corr[o + r4] = cv[b][r4];
                            args.ts[o + r4] = tv[b][r4];
                            args.pairs[o + r4] = fp;
                        }
                    }
                }
            }
        }
        if (fastm) {
            __syncthreads();
            tr_store(args.corr, a);
            __syncthreads();
#pragma unroll
            for (int rq = 0; rq < 4; ++rq)
#pragma unroll
                for (int b = 0; b < 2; ++b) tr_put(rq, b, tva[rq][b]);
            __syncthreads();
            tr_store(args.ts, a);
        }
    }
    if (fastm) {
        // pairs is one value: the wave's mirrored 64 x 64 block as rows
#pragma unroll
        for (int it = 0; it < 16; ++it) {
            const int ly = 4 * it + (lane >> 4), lx = (lane & 15) * 4;
            *reinterpret_cast<float4 *>(args.pairs + (y0 + wc * 64 + ly) * ldo + x0 + wr * 64 + lx) =
                make_float4(fp, fp, fp, fp);
        }
    }
}

// ---------------------------------------------------------------------------
// VALU tile of use = "pairwise.complete.obs": two passes per pair.
// ---------------------------------------------------------------------------
constexpr int WBM = 64;        // variables per tile side
constexpr int WBK = 32;        // samples per LDS chunk
constexpr int WTHREADS = 256;  // 16 x 16 lanes, 4 x 4 pairs each
constexpr int WLD = WBM + 4;   // LDS row length (16-byte aligned rows)

struct PairwiseArgs {
    const float *a;
    int64_t na, lda;
    const float *b;
    int64_t nb, ldb;
    int64_t dim;
    float *corr, *ts, *pairs;
    int64_t ldo;
    int64_t tiles_b;
    int64_t tile_offset, tiles_this;
    bool vec;   // ldo and the output pointers allow 16-byte stores
};

// A sample of a pair counts when both values are present (not NaN): counts
// and sums (gpuMeans, correlation.cu:7-51) give the pair's means, then the
// centred moments over the same samples (gpuSD and gpuPMCC,
// correlation.cu:53-150).  Masked samples are selected away, never multiplied
// by zero, so an Inf beside a missing value does not leak into the sums.
// Each chunk's sums are accumulated separately and then added to the totals.
template <bool SYM>
__global__ __launch_bounds__(WTHREADS) void pmcc_pairwise_kernel(PairwiseArgs args)
{
    __shared__ float lds[2][WBK][WLD];   // [0] x side, [1] y side; sample-major
    const int tid = threadIdx.x;
    const int tx = tid & 15, ty = tid >> 4;   // y = 4 tx + j, x = 4 ty + i
    const int64_t t = args.tile_offset + xcd_remap(blockIdx.x, args.tiles_this);
    int64_t I, J;
    if (SYM) {
        int ti, tj;
        tri_tile(t, ti, tj);
        I = ti;
        J = tj;
    } else {
        I = t / args.tiles_b;
        J = t % args.tiles_b;
    }
    const int64_t x0 = I * WBM, y0 = J * WBM;
    const int64_t dim = args.dim;

    // staging: variable tid >> 2, samples 8 (tid & 3) .. + 7 of the chunk;
    // samples past dim and variables past the end read as missing
    const int sv = tid >> 2, sk = (tid & 3) * 8;
    const bool va_ok = x0 + sv < args.na, vb_ok = y0 + sv < args.nb;
    const float *pa = args.a + (va_ok ? (x0 + sv) * args.lda : 0);
    const float *pb = args.b + (vb_ok ? (y0 + sv) * args.ldb : 0);
    const float qnan = __builtin_nanf("");
    float ra[8], rb[8];
    auto gload = [&](int64_t kb) {
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int64_t kk = kb + sk + q;
            ra[q] = (va_ok && kk < dim) ? pa[kk] : qnan;
            rb[q] = (vb_ok && kk < dim) ? pb[kk] : qnan;
        }
    };
    auto stage = [&]() {
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            lds[0][sk + q][sv] = ra[q];
            lds[1][sk + q][sv] = rb[q];
        }
    };

    float cnt[4][4], qa[4][4], qb[4][4], qab[4][4];

    // ---- one pass ---------------------------------------------------------
    // Each variable is shifted by c, the mean of its present values in the
    // first chunk, and the pair's count and its five shifted sums come from
    // one pass: n = sum m_a m_b, Sa = sum a' m_b, Sb = sum m_a b',
    // Saa = sum a'^2 m_b, Sbb = sum m_a b'^2, Sab = sum a' b' (a' = a - c where
    // present, else 0).  The centred moments are then Saa - Sa^2 / n, ... .
    // Where that subtraction cancels (Caa <= Saa / 16: a subset mean far from
    // c, a constant subset, fewer than two samples) or a sum is not finite
    // (Inf in the data), the tile runs the two passes below instead.
    {
        __shared__ float lds1[3][2][WBK][WLD];   // a', a'^2, presence; [side]
        gload(0);
        float csum[2] = {0.0f, 0.0f}, ccnt[2] = {0.0f, 0.0f};
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            if (fabsf(ra[q]) < INFINITY) {
                csum[0] += ra[q];
                ccnt[0] += 1.0f;
            }
            if (fabsf(rb[q]) < INFINITY) {
                csum[1] += rb[q];
                ccnt[1] += 1.0f;
            }
        }
        // the four lanes of a variable are consecutive lanes of one wave
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            csum[h] += __shfl_xor(csum[h], 1, 64);
            ccnt[h] += __shfl_xor(ccnt[h], 1, 64);
            csum[h] += __shfl_xor(csum[h], 2, 64);
            ccnt[h] += __shfl_xor(ccnt[h], 2, 64);
        }
        const float ca = ccnt[0] > 0.0f ? csum[0] / ccnt[0] : 0.0f;
        const float cb = ccnt[1] > 0.0f ? csum[1] / ccnt[1] : 0.0f;

        float sa[4][4], sb[4][4], saa[4][4], sbb[4][4], sab[4][4];
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j)
                cnt[i][j] = sa[i][j] = sb[i][j] = saa[i][j] = sbb[i][j] = sab[i][j] = 0.0f;
        for (int64_t kb = 0; kb < dim; kb += WBK) {
            __syncthreads();
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const bool pa_ = ra[q] == ra[q], pb_ = rb[q] == rb[q];
                const float da = pa_ ? ra[q] - ca : 0.0f, db = pb_ ? rb[q] - cb : 0.0f;
                lds1[0][0][sk + q][sv] = da;
                lds1[1][0][sk + q][sv] = da * da;
                lds1[2][0][sk + q][sv] = pa_ ? 1.0f : 0.0f;
                lds1[0][1][sk + q][sv] = db;
                lds1[1][1][sk + q][sv] = db * db;
                lds1[2][1][sk + q][sv] = pb_ ? 1.0f : 0.0f;
            }
            __syncthreads();
            if (kb + WBK < dim) gload(kb + WBK);
            // the second-moment sums are two-level (chunk partials, then
            // totals); Sa and Sb, small sums of shifted values that only
            // enter as Sa^2 / n, accumulate directly
            float p_aa[4][4], p_bb[4][4], p_ab[4][4];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) p_aa[i][j] = p_bb[i][j] = p_ab[i][j] = 0.0f;
#pragma unroll 2
            for (int kc = 0; kc < WBK; ++kc) {
                const float4 av4 = *reinterpret_cast<const float4 *>(&lds1[0][0][kc][ty * 4]);
                const float4 aq4 = *reinterpret_cast<const float4 *>(&lds1[1][0][kc][ty * 4]);
                const float4 am4 = *reinterpret_cast<const float4 *>(&lds1[2][0][kc][ty * 4]);
                const float4 bv4 = *reinterpret_cast<const float4 *>(&lds1[0][1][kc][tx * 4]);
                const float4 bq4 = *reinterpret_cast<const float4 *>(&lds1[1][1][kc][tx * 4]);
                const float4 bm4 = *reinterpret_cast<const float4 *>(&lds1[2][1][kc][tx * 4]);
                const float av[4] = {av4.x, av4.y, av4.z, av4.w}, aq[4] = {aq4.x, aq4.y, aq4.z, aq4.w};
                const float am[4] = {am4.x, am4.y, am4.z, am4.w};
                const float bv[4] = {bv4.x, bv4.y, bv4.z, bv4.w}, bq[4] = {bq4.x, bq4.y, bq4.z, bq4.w};
                const float bm[4] = {bm4.x, bm4.y, bm4.z, bm4.w};
#pragma unroll
                for (int i = 0; i < 4; ++i)
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        cnt[i][j] = fmaf(am[i], bm[j], cnt[i][j]);   // exact: small integers
                        sa[i][j] = fmaf(av[i], bm[j], sa[i][j]);
                        sb[i][j] = fmaf(am[i], bv[j], sb[i][j]);
                        p_aa[i][j] = fmaf(aq[i], bm[j], p_aa[i][j]);
                        p_bb[i][j] = fmaf(am[i], bq[j], p_bb[i][j]);
                        p_ab[i][j] = fmaf(av[i], bv[j], p_ab[i][j]);
                    }
            }
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    saa[i][j] += p_aa[i][j];
                    sbb[i][j] += p_bb[i][j];
                    sab[i][j] += p_ab[i][j];
                }
        }
        bool ill = false;
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const float n = cnt[i][j];
                // Sa Sb / n with the product first, so (x, y) and (y, x) round alike
                qa[i][j] = saa[i][j] - sa[i][j] * sa[i][j] / n;
                qb[i][j] = sbb[i][j] - sb[i][j] * sb[i][j] / n;
                qab[i][j] = sab[i][j] - sa[i][j] * sb[i][j] / n;
                const bool inside = x0 + ty * 4 + i < args.na && y0 + tx * 4 + j < args.nb;
                ill = ill || (inside && !(qa[i][j] > saa[i][j] * 0.0625f && qb[i][j] > sbb[i][j] * 0.0625f));
            }
        if (__syncthreads_or(ill) == 0) goto finish;
    }

    // ---- two passes (tiles the one pass does not settle) -------------------
    {
        float s1[4][4], s2[4][4];
    #pragma unroll
        for (int i = 0; i < 4; ++i)
    #pragma unroll
            for (int j = 0; j < 4; ++j) cnt[i][j] = s1[i][j] = s2[i][j] = 0.0f;

        // pass 1: count, sum x, sum y over the pair's samples.  A chunk without
        // Inf uses products with 0/1 masks (x * 0 = 0 for finite x): one fma per
        // sum instead of a select and an add, with the same sums bit for bit; a
        // chunk with an Inf selects.
        gload(0);
        for (int64_t kb = 0; kb < dim; kb += WBK) {
            __syncthreads();
            stage();
            bool inf_here = false;
    #pragma unroll
            for (int q = 0; q < 8; ++q) inf_here = inf_here || __builtin_isinf(ra[q]) || __builtin_isinf(rb[q]);
            const bool any_inf = __syncthreads_or(inf_here) != 0;
            if (kb + WBK < dim) gload(kb + WBK);
            float p1[4][4], p2[4][4];
    #pragma unroll
            for (int i = 0; i < 4; ++i)
    #pragma unroll
                for (int j = 0; j < 4; ++j) p1[i][j] = p2[i][j] = 0.0f;
            if (!any_inf) {
    #pragma unroll 4
                for (int kc = 0; kc < WBK; ++kc) {
                    const float4 a4 = *reinterpret_cast<const float4 *>(&lds[0][kc][ty * 4]);
                    const float4 b4 = *reinterpret_cast<const float4 *>(&lds[1][kc][tx * 4]);
                    const float av[4] = {a4.x, a4.y, a4.z, a4.w};
                    const float bv[4] = {b4.x, b4.y, b4.z, b4.w};
                    float am[4], a0[4], bm[4], b0[4];
    #pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        am[i] = av[i] == av[i] ? 1.0f : 0.0f;
                        a0[i] = av[i] == av[i] ? av[i] : 0.0f;
                        bm[i] = bv[i] == bv[i] ? 1.0f : 0.0f;
                        b0[i] = bv[i] == bv[i] ? bv[i] : 0.0f;
                    }
    #pragma unroll
                    for (int i = 0; i < 4; ++i)
    #pragma unroll
                        for (int j = 0; j < 4; ++j) {
                            cnt[i][j] = fmaf(am[i], bm[j], cnt[i][j]);
                            p1[i][j] = fmaf(a0[i], bm[j], p1[i][j]);
                            p2[i][j] = fmaf(b0[j], am[i], p2[i][j]);
                        }
                }
            } else {
    #pragma unroll 4
                for (int kc = 0; kc < WBK; ++kc) {
                    const float4 a4 = *reinterpret_cast<const float4 *>(&lds[0][kc][ty * 4]);
                    const float4 b4 = *reinterpret_cast<const float4 *>(&lds[1][kc][tx * 4]);
                    const float av[4] = {a4.x, a4.y, a4.z, a4.w};
                    const float bv[4] = {b4.x, b4.y, b4.z, b4.w};
    #pragma unroll
                    for (int i = 0; i < 4; ++i)
    #pragma unroll
                        for (int j = 0; j < 4; ++j) {
                            const bool m = (av[i] == av[i]) && (bv[j] == bv[j]);
                            cnt[i][j] += m ? 1.0f : 0.0f;
                            p1[i][j] += m ? av[i] : 0.0f;
                            p2[i][j] += m ? bv[j] : 0.0f;
                        }
                }
            }
    #pragma unroll
            for (int i = 0; i < 4; ++i)
    #pragma unroll
                for (int j = 0; j < 4; ++j) {
                    s1[i][j] += p1[i][j];
                    s2[i][j] += p2[i][j];
                }
        }
        // the pair's means (0 / 0 = NaN for a pair without samples: its result
        // is set below)
    #pragma unroll
        for (int i = 0; i < 4; ++i)
    #pragma unroll
            for (int j = 0; j < 4; ++j) {
                s1[i][j] = s1[i][j] / cnt[i][j];
                s2[i][j] = s2[i][j] / cnt[i][j];
            }

        // pass 2: sum dx^2, sum dy^2, sum dx dy
    #pragma unroll
        for (int i = 0; i < 4; ++i)
    #pragma unroll
            for (int j = 0; j < 4; ++j) qa[i][j] = qb[i][j] = qab[i][j] = 0.0f;
        gload(0);
        for (int64_t kb = 0; kb < dim; kb += WBK) {
            __syncthreads();
            stage();
            __syncthreads();
            if (kb + WBK < dim) gload(kb + WBK);
            float pa2[4][4], pb2[4][4], pab[4][4];
    #pragma unroll
            for (int i = 0; i < 4; ++i)
    #pragma unroll
                for (int j = 0; j < 4; ++j) pa2[i][j] = pb2[i][j] = pab[i][j] = 0.0f;
    #pragma unroll 2
            for (int kc = 0; kc < WBK; ++kc) {
                const float4 a4 = *reinterpret_cast<const float4 *>(&lds[0][kc][ty * 4]);
                const float4 b4 = *reinterpret_cast<const float4 *>(&lds[1][kc][tx * 4]);
                const float av[4] = {a4.x, a4.y, a4.z, a4.w};
                const float bv[4] = {b4.x, b4.y, b4.z, b4.w};
    #pragma unroll
                for (int i = 0; i < 4; ++i)
    #pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        const bool m = (av[i] == av[i]) && (bv[j] == bv[j]);
                        const float d = m ? av[i] - s1[i][j] : 0.0f;
                        const float e = m ? bv[j] - s2[i][j] : 0.0f;
                        pa2[i][j] = fmaf(d, d, pa2[i][j]);
                        pb2[i][j] = fmaf(e, e, pb2[i][j]);
                        pab[i][j] = fmaf(d, e, pab[i][j]);
                    }
            }
    #pragma unroll
            for (int i = 0; i < 4; ++i)
    #pragma unroll
                for (int j = 0; j < 4; ++j) {
                    qa[i][j] += pa2[i][j];
                    qb[i][j] += pb2[i][j];
                    qab[i][j] += pab[i][j];
                }
        }
    }

finish:
    // epilogue: no samples gives the reference's 0 / (0 - 1) = -0
    const bool mirror = SYM && (I != J);
    float rv[4][4], tv[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const float c = cnt[i][j];
            rv[i][j] = (c == 0.0f) ? -0.0f : qab[i][j] / (sqrtf(qa[i][j]) * sqrtf(qb[i][j]));
            tv[i][j] = t_score<true>(c, rv[i][j]);
        }
    // a lane's 4 x 4 pairs are four consecutive y per x (and four
    // consecutive x per y): 16-byte stores where the rows allow them
    const bool vec = args.vec && x0 + WBM <= args.na && y0 + WBM <= args.nb;
    if (vec) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int64_t o = (x0 + ty * 4 + i) * args.ldo + y0 + tx * 4;
            *reinterpret_cast<float4 *>(args.corr + o) = make_float4(rv[i][0], rv[i][1], rv[i][2], rv[i][3]);
            *reinterpret_cast<float4 *>(args.ts + o) = make_float4(tv[i][0], tv[i][1], tv[i][2], tv[i][3]);
            *reinterpret_cast<float4 *>(args.pairs + o) =
                make_float4(cnt[i][0], cnt[i][1], cnt[i][2], cnt[i][3]);
        }
        if (mirror) {
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int64_t o = (y0 + tx * 4 + j) * args.ldo + x0 + ty * 4;
                *reinterpret_cast<float4 *>(args.corr + o) = make_float4(rv[0][j], rv[1][j], rv[2][j], rv[3][j]);
                *reinterpret_cast<float4 *>(args.ts + o) = make_float4(tv[0][j], tv[1][j], tv[2][j], tv[3][j]);
                *reinterpret_cast<float4 *>(args.pairs + o) =
                    make_float4(cnt[0][j], cnt[1][j], cnt[2][j], cnt[3][j]);
            }
        }
        return;
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int64_t x = x0 + ty * 4 + i;
        if (x >= args.na) continue;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int64_t y = y0 + tx * 4 + j;
            if (y >= args.nb) continue;
            const int64_t o = x * args.ldo + y;
            args.corr[o] = rv[i][j];
            args.ts[o] = tv[i][j];
            args.pairs[o] = cnt[i][j];
            if (mirror) {
                const int64_t om = y * args.ldo + x;
                args.corr[om] = rv[i][j];
                args.ts[om] = tv[i][j];
                args.pairs[om] = cnt[i][j];
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Elementwise kernels
// ---------------------------------------------------------------------------

// t-scores and counts beside a correlation matrix the FP32 dot-product tiles
// wrote (use = "everything": every pair has fpairs samples).
__global__ void pmcc_finish_kernel(const float *__restrict__ corr, float *__restrict__ ts,
                                   float *__restrict__ pairs, int64_t na, int64_t nb, int64_t ldo,
                                   float fpairs)
{
    const int64_t total = na * nb;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t o = (e / nb) * ldo + e % nb;
        ts[o] = t_score<true>(fpairs, corr[o]);
        pairs[o] = fpairs;
    }
}

// dim == 0: no samples, corr = 0 / (0 - 1) = -0 for every pair.
__global__ void pmcc_empty_kernel(float *__restrict__ corr, float *__restrict__ ts,
                                  float *__restrict__ pairs, int64_t na, int64_t nb, int64_t ldo)
{
    const int64_t total = na * nb;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t o = (e / nb) * ldo + e % nb;
        corr[o] = -0.0f;
        ts[o] = t_score<true>(0.0f, -0.0f);
        pairs[o] = 0.0f;
    }
}

// gpuSignif (inst/cuda/correlation.cu:292-310): t-scores without the cap.
__global__ void ttest_kernel(const float *__restrict__ pairs, const float *__restrict__ corr,
                             float *__restrict__ ts, int64_t n)
{
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
         e += (int64_t)gridDim.x * blockDim.x)
        ts[e] = t_score<false>(pairs[e], corr[e]);
}

unsigned grid_for(int64_t total, int threads)
{
    int64_t blocks = (total + threads - 1) / threads;
    if (blocks > 65536) blocks = 65536;
    if (blocks < 1) blocks = 1;
    return (unsigned)blocks;
}

size_t align256(size_t v) { return (v + 255) / 256 * 256; }

int64_t padded_dim(int64_t dim) { return (dim + MKPAD - 1) / MKPAD * MKPAD; }

int prep_terms(int precision) { return precision == 1 ? 3 : 2; }

// Workspace layout of use = "everything": the standardised variables of a,
// then (unless same) of b.  FP32: z rows; split forms: terms, then inverse
// scales.
struct PmccWs {
    unsigned short *ta = nullptr, *tb = nullptr;
    float *isa = nullptr, *isb = nullptr, *za = nullptr, *zb = nullptr;
};

size_t side_bytes(int precision, int64_t rows, int64_t dim)
{
    if (precision == 0) return align256((size_t)rows * (size_t)dim * sizeof(float));
    return align256((size_t)prep_terms(precision) * (size_t)rows * (size_t)padded_dim(dim) *
                    sizeof(unsigned short)) +
           align256((size_t)rows * sizeof(float));
}

PmccWs carve(int precision, int64_t na, int64_t nb, bool same, int64_t dim, void *ws)
{
    PmccWs w;
    char *p = static_cast<char *>(ws);
    auto side = [&](int64_t rows, unsigned short *&t, float *&is, float *&z) {
        if (precision == 0) {
            z = reinterpret_cast<float *>(p);
        } else {
            t = reinterpret_cast<unsigned short *>(p);
            is = reinterpret_cast<float *>(
                p + align256((size_t)prep_terms(precision) * (size_t)rows *
                             (size_t)padded_dim(dim) * sizeof(unsigned short)));
        }
        p += side_bytes(precision, rows, dim);
    };
    side(na, w.ta, w.isa, w.za);
    if (same) {
        w.tb = w.ta;
        w.isb = w.isa;
        w.zb = w.za;
    } else {
        side(nb, w.tb, w.isb, w.zb);
    }
    return w;
}

hipError_t launch_prep(int precision, const float *x, int64_t n, int64_t dim, int64_t ldx,
                       unsigned short *terms, float *iscale, float *z, hipStream_t stream)
{
    int64_t blocks = (n + 3) / 4;
    if (blocks > 65536) blocks = 65536;
    const int64_t kp = padded_dim(dim);
    if (precision == 0)
        hipLaunchKernelGGL(pmcc_prep_kernel<PREP_FP32>, dim3((unsigned)blocks), dim3(256), 0,
                           stream, x, n, dim, ldx, kp, terms, iscale, z);
    else if (precision == 1)
        hipLaunchKernelGGL(pmcc_prep_kernel<PREP_BF16>, dim3((unsigned)blocks), dim3(256), 0,
                           stream, x, n, dim, ldx, kp, terms, iscale, z);
    else
        hipLaunchKernelGGL(pmcc_prep_kernel<PREP_FP16>, dim3((unsigned)blocks), dim3(256), 0,
                           stream, x, n, dim, ldx, kp, terms, iscale, z);
    return hipGetLastError();
}

template <int SPLIT, bool SYM, bool VEC>
hipError_t launch_tile(PmccTileArgs &args, int64_t ids, hipStream_t stream)
{
    const int64_t max_ids = int64_t(1) << 23;
    for (int64_t off = 0; off < ids; off += max_ids) {
        args.id_offset = off;
        args.ids_this = (ids - off) < max_ids ? (ids - off) : max_ids;
        hipLaunchKernelGGL((pmcc_tile_kernel<SPLIT, SYM, VEC>), dim3((unsigned)args.ids_this),
                           dim3(MTHREADS), 0, stream, args);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

template <int SPLIT>
hipError_t launch_tile_split(bool same, bool vec, PmccTileArgs &args, int64_t ids,
                             hipStream_t stream)
{
    if (same)
        return vec ? launch_tile<SPLIT, true, true>(args, ids, stream)
                   : launch_tile<SPLIT, true, false>(args, ids, stream);
    return launch_tile<SPLIT, false, false>(args, ids, stream);
}

}  // namespace

size_t pmcc_workspace_bytes(int use, int precision, int64_t na, int64_t nb, bool same,
                            int64_t dim)
{
    if (use != 0 || dim <= 0 || na <= 0 || nb <= 0) return 0;
    return side_bytes(precision, na, dim) + (same ? 0 : side_bytes(precision, nb, dim));
}

hipError_t launch_pmcc(int use, int precision, const float *a, int64_t na, int64_t lda,
                       const float *b, int64_t nb, int64_t ldb, bool same, int64_t dim,
                       float *pairs, float *corr, float *ts, int64_t ldo, void *workspace,
                       hipStream_t stream)
{
    if (na <= 0 || nb <= 0) return hipSuccess;
    if (dim == 0) {
        hipLaunchKernelGGL(pmcc_empty_kernel, dim3(grid_for(na * nb, 256)), dim3(256), 0, stream,
                           corr, ts, pairs, na, nb, ldo);
        return hipGetLastError();
    }
    if (use == 1) {
        PairwiseArgs args;
        args.a = a;
        args.na = na;
        args.lda = lda;
        args.b = b;
        args.nb = nb;
        args.ldb = ldb;
        args.dim = dim;
        args.corr = corr;
        args.ts = ts;
        args.pairs = pairs;
        args.ldo = ldo;
        args.vec = (ldo % 4 == 0) && ((reinterpret_cast<uintptr_t>(corr) | reinterpret_cast<uintptr_t>(ts) |
                                       reinterpret_cast<uintptr_t>(pairs)) & 15u) == 0;
        const int64_t ta = (na + WBM - 1) / WBM, tb = (nb + WBM - 1) / WBM;
        args.tiles_b = tb;
        const int64_t total = same ? ta * (ta + 1) / 2 : ta * tb;
        const int64_t max_tiles = int64_t(1) << 23;
        for (int64_t off = 0; off < total; off += max_tiles) {
            args.tile_offset = off;
            args.tiles_this = (total - off) < max_tiles ? (total - off) : max_tiles;
            if (same)
                hipLaunchKernelGGL(pmcc_pairwise_kernel<true>, dim3((unsigned)args.tiles_this),
                                   dim3(WTHREADS), 0, stream, args);
            else
                hipLaunchKernelGGL(pmcc_pairwise_kernel<false>, dim3((unsigned)args.tiles_this),
                                   dim3(WTHREADS), 0, stream, args);
            hipError_t e = hipGetLastError();
            if (e != hipSuccess) return e;
        }
        return hipSuccess;
    }

    PmccWs w = carve(precision, na, nb, same, dim, workspace);
    hipError_t e = launch_prep(precision, a, na, dim, lda, w.ta, w.isa, w.za, stream);
    if (e != hipSuccess) return e;
    if (!same) {
        e = launch_prep(precision, b, nb, dim, ldb, w.tb, w.isb, w.zb, stream);
        if (e != hipSuccess) return e;
    }
    const float fpairs = (float)dim;
    if (precision == 0) {
        // corr[x * ldo + y] = z_x . z_y: the plain cross tile with x on its
        // row side and y on its column side; with a == b the symmetric
        // layout computes the lower tiles and mirrors them
        if (same)
            e = launch_plain(OP_DOT, OUT_FULL, w.za, na, dim, w.za, na, dim, dim, 2.0f, corr, ldo,
                             stream);
        else
            e = launch_plain(OP_DOT, OUT_CROSS, w.zb, nb, dim, w.za, na, dim, dim, 2.0f, corr, ldo,
                             stream);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(pmcc_finish_kernel, dim3(grid_for(na * nb, 256)), dim3(256), 0, stream,
                           corr, ts, pairs, na, nb, ldo, fpairs);
        return hipGetLastError();
    }
    PmccTileArgs args;
    args.ta = w.ta;
    args.tb = w.tb;
    args.na = na;
    args.nb = nb;
    args.kp = padded_dim(dim);
    args.isa = w.isa;
    args.isb = w.isb;
    args.corr = corr;
    args.ts = ts;
    args.pairs = pairs;
    args.ldo = ldo;
    args.fpairs = fpairs;
    args.tiles_a = (na + MBM - 1) / MBM;
    args.tiles_b = (nb + MBM - 1) / MBM;
    const int64_t sa = (args.tiles_a + MSUPER - 1) / MSUPER;
    const int64_t sbt = (args.tiles_b + MSUPER - 1) / MSUPER;
    args.super_b = sbt;
    const int64_t ids = (same ? sa * (sa + 1) / 2 : sa * sbt) * MSUPER * MSUPER;
    const bool vec = (ldo % 4 == 0) && ((reinterpret_cast<uintptr_t>(corr) |
                                         reinterpret_cast<uintptr_t>(ts) |
                                         reinterpret_cast<uintptr_t>(pairs)) & 15u) == 0;
    return precision == 1 ? launch_tile_split<SPLIT_BF16>(same, vec, args, ids, stream)
                          : launch_tile_split<SPLIT_FP16>(same, vec, args, ids, stream);
}

hipError_t launch_ttest(const float *pairs, const float *corr, float *ts, int64_t n,
                        hipStream_t stream)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(ttest_kernel, dim3(grid_for(n, 256)), dim3(256), 0, stream, pairs, corr,
                       ts, n);
    return hipGetLastError();
}

}  // namespace gt
