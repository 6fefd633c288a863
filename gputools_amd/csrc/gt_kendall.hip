// gt_kendall.hip -- Kendall's tau of gpuCor(method = "kendall") on gfx950.
//
// Replaces masterKendall (src/kendall.cpp:13-55) and gpuKendall
// (inst/cuda/kendall.cu:4-47).  For variables x (of A) and y (of B) over n
// samples, gputools counts the sample pairs i < j that are concordant, where
//   (x_j > x_i and y_j > y_i) or (x_j < x_i and y_j < y_i)
//   or (x_j == x_i and y_j == y_i)                        (kendall.cu:18-22)
// (any comparison with NaN is false), and returns
//   (concordant - discordant) / (n (n - 1) / 2),  discordant = that - concordant
// with the final steps in FP32 and the quotient in FP64 (kendall.cu:40-45).
// It runs one 256-thread block per pair of variables, each re-comparing the
// raw values of both variables n^2/2 times.
//
// Here the comparisons are made once per variable.  A planes pass encodes
// each sample pair (i, j), i < j, of a variable as three bits: P (x_j > x_i),
// N (x_j < x_i) and E (x_j == x_i); a NaN on either side clears all three.
// A wave compares x_i with 64 consecutive x_j at once (the loads of up to 32
// such steps issued together) and __ballot turns the comparisons into bit
// words, which v_writelane collects across lanes for coalesced 256-byte
// stores.  In a variable without NaN every sample pair is exactly one of P,
// N and E, so only P and E are compared and N = ~(P | E) within the row.
// The sample pairs of row i take
// ceil((n - 1 - i) / 32) words (the tail bits are zero), so
//   concordant(x, y) = sum over words of popc(Px & Py) + popc(Nx & Ny)
//                      + popc(Ex & Ey)
// exactly, 6 VALU operations per 32 sample pairs per pair of variables (AND
// and v_bcnt_u32_b32, which adds the count to an accumulator), or 4 where the
// E words of the staged chunk are all zero (ties are rare bits in continuous
// data; a side without any tied variable does not even load them).  In a
// tile without NaN, a chunk without E bits has exactly one of the P and N
// bits of every sample pair set, so
//   concordant(x, y) = sum over words of popc(~(Px ^ Py)) - (bits past row ends),
// 2 operations per 32 sample pairs (the bits past row ends depend on the
// layout alone: a small pass tabulates them per chunk).  The count
// kernel tiles 128 x 128 pairs of variables over 256 lanes (8 x 8 each),
// stages word chunks of both sides through LDS and splits the word range over
// workgroups, adding its counts with 64-bit atomics.  A last pass applies the
// reference's FP32 / FP64 formula to the exact count.
//
// The reference accumulates the count in FP32 (kendall.cu:14, 23, 28-38):
// below 2^24 concordant pairs (always when n <= 5793) that sum is exact and
// the results are identical; above it the reference's sum can round, and this
// build keeps the exact count.
#include "gt_common.h"
#include "gt_internal.h"

#include <type_traits>

namespace gt {
namespace {

// Words of the rows with m = 1 .. M sample pairs: sum_{t=1}^{M} ceil(t / 32).
__host__ __device__ __forceinline__ int64_t words_upto(int64_t m)
{
    const int64_t q = m / 32, r = m % 32;
    return 16 * q * (q + 1) + r * (q + 1);
}

// v_writelane_b32 through the compiler's own intrinsic (this clang has no
// builtin for it), so that the compiler sees the instruction and places the
// wait states between the compare that writes the ballot's SGPRs and the
// lane write that reads them.
extern "C" __device__ int gt_amdgcn_writelane(int value, int lane, int old)
    __asm("llvm.amdgcn.writelane.i32");

// One plane's two words of one step (a 64-bit ballot) into lanes L and
// L + 1 of the plane words: v_writelane_b32 with the lane select as an
// inline constant.
template <int L>
__device__ __forceinline__ void put_plane(uint32_t &w, unsigned long long bits)
{
    w = (uint32_t)gt_amdgcn_writelane((int)(uint32_t)bits, L, (int)w);
    w = (uint32_t)gt_amdgcn_writelane((int)(uint32_t)(bits >> 32), L + 1, (int)w);
}

// The three planes' words of one step.
template <int L>
__device__ __forceinline__ void put_step(uint32_t &wp, uint32_t &wn, uint32_t &we,
                                         unsigned long long gt, unsigned long long lt,
                                         unsigned long long eq)
{
    put_plane<L>(wp, gt);
    put_plane<L>(wn, lt);
    put_plane<L>(we, eq);
}

// f(std::integral_constant<int, U>) for U = 0 .. N - 1, in order.
template <int U, int N, typename F>
__device__ __forceinline__ void static_for(F &&f)
{
    if constexpr (U < N) {
        f(std::integral_constant<int, U>{});
        static_for<U + 1, N>(f);
    }
}

// ---------------------------------------------------------------------------
// Planes pass: one workgroup per (variable, row split), one wave per sample i.
// ---------------------------------------------------------------------------
// Plane p (0 = P, 1 = N, 2 = E) of variable v, word w at
// planes[(p * nvar + v) * wpad + w]; row i starts at word
// words_upto(n - 1) - words_upto(n - 1 - i).  Words past the last row are
// zero (written by this pass in P and N; the E planes are cleared before it,
// since E words are written only where a step has a tie).  flags[v] (cleared before) gets bit 0 when
// any sample pair of v is tied (its E plane has a set bit) and bit 1 when v
// holds a NaN (a block-wide vote over all n values).
template <bool LX>
__global__ __launch_bounds__(256) void kendall_planes_kernel(const float *__restrict__ x,
                                                             int64_t nvar, int64_t n, int64_t ldx,
                                                             uint32_t *__restrict__ planes,
                                                             int64_t wpad,
                                                             uint32_t *__restrict__ ties,
                                                             int rsplit)
{
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int64_t wtot = words_upto(n - 1);
    const float qnan = __builtin_nanf("");
    const int64_t pstride = nvar * wpad;
    // workgroup (variable v, split rs); its waves take rows
    // i = 4 rs + wave + 4 rsplit k, a mix of long and short rows
    const int64_t v = blockIdx.x / rsplit;
    const int rs = (int)(blockIdx.x % rsplit);
    const float *xv = x + v * ldx;
    uint32_t *const pv = planes + v * wpad;
    // Does the variable hold a NaN?  Without one, every sample pair in a row
    // is exactly one of P, N and E, so the N word is ~(P | E) within the row
    // and needs no compare of its own; the E words need lane writes only in
    // steps with a tie.  That takes a step from 3 compares and 6 lane writes
    // to 2 and 2.
    // LX: the variable's n values are staged in LDS (the same pass), followed
    // by 2048 NaNs, and the rows read them from there instead of through the
    // vector memory path.  A group reads all 32 steps, up to x_j with
    // j = n - 1 + 63 + 31 * 64, so the NaN tail stands in for the bounds and
    // step tests (a NaN compares false).
    extern __shared__ float xs[];
    bool nan_here = false;
    for (int64_t t = threadIdx.x; t < n; t += blockDim.x) {
        const float val = xv[t];
        nan_here = nan_here || val != val;
        if (LX) xs[t] = val;
    }
    if (LX)
        for (int t = threadIdx.x; t < 2048; t += blockDim.x) xs[n + t] = qnan;
    const bool has_nan = __syncthreads_or(nan_here) != 0;   // (and the LDS copy is complete)
    // P and N are written in full, up to wpad: split 0 zeroes the words after
    // the last row (fewer than KKW); the E planes are cleared before this pass
    if (rs == 0 && threadIdx.x < wpad - wtot) {
        pv[wtot + threadIdx.x] = 0u;
        pv[pstride + wtot + threadIdx.x] = 0u;
    }
    bool tied = false;
    for (int64_t i = 4 * rs + wave; i < n - 1; i += 4 * (int64_t)rsplit) {
        const int64_t m = n - 1 - i;
        const float xi = LX ? xs[i] : xv[i];
        const float *xr = xv + i + 1;                    // x_j, j = i + 1 + q
        uint32_t *const row = pv + (wtot - words_upto(m));
        const int64_t wr = (m + 31) / 32;
        const int64_t steps = (m + 63) / 64;
        // groups of 32 steps: lanes 2u and 2u + 1 collect the words of step
        // u (a compile-time lane, so each word is one v_writelane), and the
        // group's 64 words per plane go out in one store per plane
        for (int64_t g0 = 0; g0 < steps; g0 += 32) {
            int nst = (int)((steps - g0) < 32 ? (steps - g0) : 32);
            // a 32-bit scalar: the step tests below stay scalar compares
            // (the compiler otherwise tests the 64-bit value on the VALU)
            asm("" : "+s"(nst));
            float xj[32];
            // every xj[u] a register of its own (compile-time u): a loop with
            // a run-time exit would index xj dynamically and wait on each load
            if (LX) {
                // n <= 8192: 32-bit offsets; step u is an immediate offset
                const float *xl = xs + (int)(i + 1 + g0 * 64) + lane;
                static_for<0, 32>([&](auto uc) {
                    constexpr int u = decltype(uc)::value;
                    xj[u] = xl[u * 64];
                });
            } else {
                static_for<0, 32>([&](auto uc) {
                    constexpr int u = decltype(uc)::value;
                    if (u < nst) {
                        const int64_t q = (g0 + u) * 64 + lane;
                        xj[u] = q < m ? xr[q] : qnan;
                    }
                });
            }
            uint32_t wp = 0, wn = 0, we = 0;
            bool group_tied = false;
            const int64_t w = 2 * g0 + lane;
            if (has_nan) {
                static_for<0, 32>([&](auto uc) {
                    constexpr int u = decltype(uc)::value;
                    if (u < nst) {
                        const unsigned long long gt = __builtin_amdgcn_ballot_w64(xj[u] > xi);
                        const unsigned long long lt = __builtin_amdgcn_ballot_w64(xj[u] < xi);
                        const unsigned long long eq = __builtin_amdgcn_ballot_w64(xj[u] == xi);
                        put_step<2 * u>(wp, wn, we, gt, lt, eq);
                        group_tied = group_tied || eq != 0ull;
                    }
                });
            } else {
                static_for<0, 32>([&](auto uc) {
                    constexpr int u = decltype(uc)::value;
                    if (u < nst) {
                        const unsigned long long gt = __builtin_amdgcn_ballot_w64(xj[u] > xi);
                        const unsigned long long eq = __builtin_amdgcn_ballot_w64(xj[u] == xi);
                        put_plane<2 * u>(wp, gt);
                        if (eq != 0ull) {
                            put_plane<2 * u>(we, eq);
                            group_tied = true;
                        }
                    }
                });
                // the sample pairs of word w are q = 32 w .. 32 w + 31 < m
                const int64_t rem = m - 32 * w;
                const uint32_t valid = rem >= 32 ? 0xffffffffu : rem > 0 ? (1u << rem) - 1u : 0u;
                wn = ~(wp | we) & valid;
            }
            // lanes at or past `used` hold no word of this group; w < wr
            // drops the odd last word of the row (it belongs to the next row)
            const int used = 2 * nst;
            if (lane < used && w < wr) {
                row[w] = wp;
                row[pstride + w] = wn;
                // E words that are zero stay as cleared
                if (group_tied) row[2 * pstride + w] = we;
            }
            tied = tied || group_tied;
        }
    }
    const uint32_t bits = (tied ? 1u : 0u) | (has_nan ? 2u : 0u);
    if (lane == 0 && bits) atomicOr(ties + v, bits);
}

// ---------------------------------------------------------------------------
// Count pass: popcount tiles.
// ---------------------------------------------------------------------------
constexpr int KBM = 128;                 // variables per tile side
constexpr int KKW = 8;                   // words per LDS chunk (256 sample pairs)
constexpr int KTHREADS = 256;            // 16 x 16 lanes, 8 x 8 pairs each
constexpr int KUNITS = 2 * 3 * KBM * 2 / KTHREADS;  // 16-byte staging units per lane

// Bits past row ends per chunk of KKW words, from the layout alone: the last
// word of row i (m = n - 1 - i sample pairs) holds m % 32 of them when that is
// not 0, and the words after the last row (up to wpad) are past the end
// entirely.  pe (cleared before) gets one count per chunk.
__global__ __launch_bounds__(256) void kendall_pastend_kernel(int64_t n, int64_t wpad,
                                                              uint32_t *__restrict__ pe)
{
    const int64_t wtot = words_upto(n - 1);
    const int64_t rows = n - 1, tail = wpad - wtot;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < rows + tail;
         t += (int64_t)gridDim.x * blockDim.x) {
        if (t < rows) {
            const int64_t m = n - 1 - t;
            if (m % 32 != 0) {
                const int64_t last = wtot - words_upto(m) + (m + 31) / 32 - 1;
                atomicAdd(pe + last / KKW, 32u - (uint32_t)(m % 32));
            }
        } else {
            atomicAdd(pe + (wtot + (t - rows)) / KKW, 32u);
        }
    }
}

struct KendallArgs {
    const uint32_t *pa, *pb;   // planes of A (na variables) and B (nb variables)
    int64_t na, nb, wpad;
    unsigned long long *count; // na x nb (row stride ldc), pair (x, y) at x * ldc + y
    int64_t ldc;
    int64_t tiles_a, tiles_b, ntiles;
    int64_t split_w;           // words per workgroup (multiple of KKW)
    const uint32_t *pe;        // per chunk of KKW words: bits past row ends
    const uint32_t *ties_a, *ties_b;   // per variable: bit 0 a tie (E bit set), bit 1 a NaN
    int64_t id_offset, ids_this;
};

// Accumulate one staged chunk into the 8 x 8 counts of this lane: per plane
// and pair one AND and one v_bcnt_u32_b32 (popcount plus accumulator); the
// empty asm keeps the compiler from regrouping the sums into an extra v_add3.
template <int NP>
__device__ __forceinline__ void count_chunk(const uint32_t (&lds)[2][3][KKW][KBM], int tx, int ty,
                                            uint32_t (&cnt)[8][8])
{
#pragma unroll 2
    for (int kw = 0; kw < KKW; ++kw) {
        uint32_t xa[NP][8], yb[NP][8];
#pragma unroll
        for (int p = 0; p < NP; ++p) {
            const uint4 a0 = *reinterpret_cast<const uint4 *>(&lds[0][p][kw][ty * 4]);
            const uint4 a1 = *reinterpret_cast<const uint4 *>(&lds[0][p][kw][64 + ty * 4]);
            const uint4 b0 = *reinterpret_cast<const uint4 *>(&lds[1][p][kw][tx * 4]);
            const uint4 b1 = *reinterpret_cast<const uint4 *>(&lds[1][p][kw][64 + tx * 4]);
            xa[p][0] = a0.x; xa[p][1] = a0.y; xa[p][2] = a0.z; xa[p][3] = a0.w;
            xa[p][4] = a1.x; xa[p][5] = a1.y; xa[p][6] = a1.z; xa[p][7] = a1.w;
            yb[p][0] = b0.x; yb[p][1] = b0.y; yb[p][2] = b0.z; yb[p][3] = b0.w;
            yb[p][4] = b1.x; yb[p][5] = b1.y; yb[p][6] = b1.z; yb[p][7] = b1.w;
        }
#pragma unroll
        for (int r = 0; r < 8; ++r)
#pragma unroll
            for (int c = 0; c < 8; ++c) {
                uint32_t s = cnt[r][c];
#pragma unroll
                for (int p = 0; p < NP; ++p) {
                    s += __builtin_popcount(xa[p][r] & yb[p][c]);
                    asm("" : "+v"(s));
                }
                cnt[r][c] = s;
            }
    }
}

// Chunks without E bits in tiles without NaN: every sample pair has exactly
// one of its P and N bits, so the pair is concordant where the P bits agree:
// one XNOR and one v_bcnt_u32_b32 per word and pair.  The bits past the end
// of a row agree too (all zero); the caller subtracts them per chunk.
__device__ __forceinline__ void count_chunk_xnor(const uint32_t (&lds)[2][3][KKW][KBM], int tx,
                                                 int ty, uint32_t (&cnt)[8][8])
{
    // fully unrolled: a rolled loop here is a loop of its own, and the
    // compiler then copies the 64 counts into and out of its registers on
    // every chunk
#pragma unroll
    for (int kw = 0; kw < KKW; ++kw) {
        const uint4 a0 = *reinterpret_cast<const uint4 *>(&lds[0][0][kw][ty * 4]);
        const uint4 a1 = *reinterpret_cast<const uint4 *>(&lds[0][0][kw][64 + ty * 4]);
        const uint4 b0 = *reinterpret_cast<const uint4 *>(&lds[1][0][kw][tx * 4]);
        const uint4 b1 = *reinterpret_cast<const uint4 *>(&lds[1][0][kw][64 + tx * 4]);
        const uint32_t xa[8] = {a0.x, a0.y, a0.z, a0.w, a1.x, a1.y, a1.z, a1.w};
        const uint32_t yb[8] = {b0.x, b0.y, b0.z, b0.w, b1.x, b1.y, b1.z, b1.w};
#pragma unroll
        for (int r = 0; r < 8; ++r)
#pragma unroll
            for (int c = 0; c < 8; ++c) {
                uint32_t s = cnt[r][c];
                s += __builtin_popcount(~(xa[r] ^ yb[c]));
                asm("" : "+v"(s));
                cnt[r][c] = s;
            }
    }
}

// SYM: pa == pb; tiles with J <= I only (the finish pass reads the other half
// from the transposed position).
template <bool SYM>
__global__ __launch_bounds__(KTHREADS) void kendall_count_kernel(KendallArgs args)
{
    __shared__ uint32_t lds[2][3][KKW][KBM];   // [side][plane][word][variable]
    const int tid = threadIdx.x;
    const int tx = tid & 15, ty = tid >> 4;    // y = tx*4 + {0..3, 64..67}, x likewise with ty
    const int64_t id = args.id_offset + xcd_remap(blockIdx.x, args.ids_this);
    const int64_t split = id / args.ntiles, t = id % args.ntiles;
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
    const int64_t x0 = I * KBM, y0 = J * KBM;
    const int64_t w_begin = split * args.split_w;
    int64_t w_end = w_begin + args.split_w;
    if (w_end > args.wpad) w_end = args.wpad;
    if (w_begin >= w_end) return;

    // Flags of the tile's variables (A in lanes 0..127, B in 128..255).  The
    // E planes only matter when a variable on each side has a tie; for
    // tie-free tiles (continuous data) they are all zero and are skipped:
    // two planes, 4 instead of 6 operations per word.  A tile with no tie and
    // no NaN at all takes the XOR form below.
    uint32_t f = 0;
    if (tid < KBM) {
        const int64_t v = x0 + tid;
        f = v < args.na ? args.ties_a[v] : 0u;
    } else {
        const int64_t v = y0 + (tid - KBM);
        f = v < args.nb ? args.ties_b[v] : 0u;
    }
    const bool tie_x = __syncthreads_or(tid < KBM && (f & 1u)) != 0;
    const bool tie_y = __syncthreads_or(tid >= KBM && (f & 1u)) != 0;
    const bool use_e = tie_x && tie_y;
    // any NaN in the tile: the XNOR form below is not used
    const bool nan_tile = __syncthreads_or((f & 2u) != 0u) != 0;

    // staging unit u = tid + 256 q: half u & 1 (words 4 half .. +3 of the
    // chunk), variable (u >> 1) & 127, side/plane u >> 8 (side = sp / 3)
    const uint32_t *src[KUNITS];
    bool ok[KUNITS], inr[KUNITS];
#pragma unroll
    for (int q = 0; q < KUNITS; ++q) {
        const int u = tid + KTHREADS * q;
        const int sp = u >> 8, v = (u >> 1) & (KBM - 1), half = u & 1;
        const int side = sp / 3, plane = sp % 3;
        const int64_t nside = side ? args.nb : args.na;
        const int64_t gv = (side ? y0 : x0) + v;
        // E words are loaded for a side with a tied variable: E & E needs
        // both sides, the XNOR form's check needs either.  N words are loaded
        // ahead only in a tile with a NaN; other chunks that need them load
        // them when their E bits show up.
        inr[q] = gv < nside;
        ok[q] = inr[q] && (plane == 0 || (plane == 1 && nan_tile) ||
                           (plane == 2 && (side ? tie_y : tie_x)));
        src[q] = (side ? args.pb : args.pa) + ((int64_t)plane * nside + (inr[q] ? gv : 0)) * args.wpad +
                 half * 4;
    }
    uint4 stage[KUNITS];
    auto gload = [&](int64_t w) {
#pragma unroll
        for (int q = 0; q < KUNITS; ++q)
            stage[q] = ok[q] ? *reinterpret_cast<const uint4 *>(src[q] + w) : make_uint4(0, 0, 0, 0);
    };

    uint32_t cnt[8][8];
#pragma unroll
    for (int r = 0; r < 8; ++r)
#pragma unroll
        for (int c = 0; c < 8; ++c) cnt[r][c] = 0;

    uint32_t past_end = 0;   // XNOR chunks: bits past row ends, counted for every pair
    gload(w_begin);
    for (int64_t w = w_begin; w < w_end; w += KKW) {
        __syncthreads();
#pragma unroll
        for (int q = 0; q < KUNITS; ++q) {
            const int u = tid + KTHREADS * q;
            const int sp = u >> 8, v = (u >> 1) & (KBM - 1), half = u & 1;
            uint32_t *d = &lds[sp / 3][sp % 3][half * 4][v];
            d[0 * KBM] = stage[q].x;
            d[1 * KBM] = stage[q].y;
            d[2 * KBM] = stage[q].z;
            d[3 * KBM] = stage[q].w;
        }
        // units 2 and 5 of every lane are the E words of the two sides: a
        // chunk where none is set (ties are rare bits) counts two planes
        const uint4 e0 = stage[2], e1 = stage[5];
        const bool e_here = (e0.x | e0.y | e0.z | e0.w | e1.x | e1.y | e1.z | e1.w) != 0u;
        const bool any_e = __syncthreads_or(e_here) != 0;
        if (any_e && !nan_tile) {
            // this chunk counts P and N: load its N words (units 1 and 4)
            uint4 sn[2];
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int q = 1 + 3 * h;
                sn[h] = inr[q] ? *reinterpret_cast<const uint4 *>(src[q] + w) : make_uint4(0, 0, 0, 0);
            }
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int u = tid + KTHREADS * (1 + 3 * h);
                const int sp = u >> 8, v = (u >> 1) & (KBM - 1), half = u & 1;
                uint32_t *d = &lds[sp / 3][sp % 3][half * 4][v];
                d[0 * KBM] = sn[h].x;
                d[1 * KBM] = sn[h].y;
                d[2 * KBM] = sn[h].z;
                d[3 * KBM] = sn[h].w;
            }
            __syncthreads();
        }
        if (w + KKW < w_end) gload(w + KKW);
        if (!any_e && !nan_tile) {
            // the chunk's bits past row ends (the same for every variable)
            past_end += args.pe[w / KKW];
            count_chunk_xnor(lds, tx, ty, cnt);
        } else if (any_e && use_e) {
            count_chunk<3>(lds, tx, ty, cnt);
        } else {
            count_chunk<2>(lds, tx, ty, cnt);
        }
    }

#pragma unroll
    for (int r = 0; r < 8; ++r) {
        const int64_t x = x0 + (r < 4 ? ty * 4 + r : 64 + ty * 4 + (r - 4));
        if (x >= args.na) continue;
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            const int64_t y = y0 + (c < 4 ? tx * 4 + c : 64 + tx * 4 + (c - 4));
            if (y >= args.nb) continue;
            atomicAdd(args.count + x * args.ldc + y, (unsigned long long)(cnt[r][c] - past_end));
        }
    }
}

// tau from the exact count, with the reference's arithmetic (kendall.cu:40-45);
// count == nullptr: no sample pairs (n < 2), the count is 0.
__global__ void kendall_finish_kernel(const unsigned long long *__restrict__ count, int64_t ldc,
                                      int64_t na, int64_t nb, int64_t n, int sym,
                                      double *__restrict__ tau, int64_t ldt)
{
    const int64_t total = na * nb;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t x = e % na, y = e / na;
        const bool lower = !sym || (x / KBM >= y / KBM);
        const unsigned long long c = count ? (lower ? count[x * ldc + y] : count[y * ldc + x]) : 0ull;
        const float concordant = (float)c;
        float denom = (float)n;
        denom = (denom * (denom - 1.0f)) / 2.0f;
        const float discordant = denom - concordant;
        const float numer = concordant - discordant;
        tau[y * ldt + x] = (double)numer / (double)denom;
    }
}

unsigned grid_for(int64_t total, int threads)
{
    int64_t blocks = (total + threads - 1) / threads;
    if (blocks > 65536) blocks = 65536;
    if (blocks < 1) blocks = 1;
    return (unsigned)blocks;
}

size_t align256(size_t v) { return (v + 255) / 256 * 256; }

int64_t padded_words(int64_t n)
{
    const int64_t w = n > 1 ? words_upto(n - 1) : 0;
    return (w + KKW - 1) / KKW * KKW;
}

}  // namespace

size_t kendall_workspace_bytes(int64_t na, int64_t nb, bool same, int64_t n)
{
    if (na <= 0 || nb <= 0 || n < 2) return 0;
    const size_t wpad = (size_t)padded_words(n);
    const size_t planes = align256(3 * (size_t)na * wpad * 4) +
                          (same ? 0 : align256(3 * (size_t)nb * wpad * 4));
    const size_t ties = align256((size_t)na * 4) + (same ? 0 : align256((size_t)nb * 4));
    const size_t pe = align256(wpad / KKW * 4);
    return planes + ties + pe + align256((size_t)na * (size_t)nb * 8);
}

hipError_t launch_kendall(const float *a, int64_t na, int64_t lda, const float *b, int64_t nb,
                          int64_t ldb, bool same, int64_t n, double *tau, int64_t ldt,
                          void *workspace, hipStream_t stream)
{
    if (na <= 0 || nb <= 0) return hipSuccess;
    hipError_t e;
    const int64_t wpad = padded_words(n);
    char *ws = static_cast<char *>(workspace);
    uint32_t *pa = nullptr, *pb = nullptr, *ta = nullptr, *tb = nullptr, *pe = nullptr;
    unsigned long long *count = nullptr;
    if (n >= 2) {
        pa = reinterpret_cast<uint32_t *>(ws);
        ws += align256(3 * (size_t)na * (size_t)wpad * 4);
        pb = pa;
        if (!same) {
            pb = reinterpret_cast<uint32_t *>(ws);
            ws += align256(3 * (size_t)nb * (size_t)wpad * 4);
        }
        ta = reinterpret_cast<uint32_t *>(ws);
        ws += align256((size_t)na * 4);
        tb = ta;
        if (!same) {
            tb = reinterpret_cast<uint32_t *>(ws);
            ws += align256((size_t)nb * 4);
        }
        pe = reinterpret_cast<uint32_t *>(ws);
        ws += align256((size_t)(wpad / KKW) * 4);
        count = reinterpret_cast<unsigned long long *>(ws);
        // zero the E planes (written only where a step has a tie), tie flags,
        // past-end counts and counts with fill kernels, not memset nodes, so
        // that a captured graph replays correctly; the planes pass writes P
        // and N in full
        e = launch_fill(reinterpret_cast<float *>(pa + 2 * (size_t)na * (size_t)wpad),
                        na * wpad, 0.0f, stream);
        if (e != hipSuccess) return e;
        if (!same) {
            e = launch_fill(reinterpret_cast<float *>(pb + 2 * (size_t)nb * (size_t)wpad),
                            nb * wpad, 0.0f, stream);
            if (e != hipSuccess) return e;
        }
        const size_t zero_floats =
            (size_t)(reinterpret_cast<char *>(count + (size_t)na * (size_t)nb) -
                     reinterpret_cast<char *>(ta)) / 4;
        e = launch_fill(reinterpret_cast<float *>(ta), (int64_t)zero_floats, 0.0f, stream);
        if (e != hipSuccess) return e;
        {
            // row splits per variable: about 16 rows per wave, within the
            // grid limit (every row is still covered: waves stride over rows)
            int64_t rs64 = (n + 63) / 64;
            const int64_t nv = na > nb ? na : nb;
            if (nv * rs64 > 0x7fffffff) rs64 = 0x7fffffff / nv;
            const int rsplit = (int)(rs64 > 0 ? rs64 : 1);
            // up to 8192 samples (32 KB + an 8 KB tail) the rows read the variable from LDS
            const bool lx = n <= 8192;
            const unsigned lds = lx ? (unsigned)((n + 2048) * 4) : 0u;
            auto planes_fn = lx ? kendall_planes_kernel<true> : kendall_planes_kernel<false>;
            hipLaunchKernelGGL(planes_fn, dim3((unsigned)(na * rsplit)), dim3(256), lds, stream,
                               a, na, n, lda, pa, wpad, ta, rsplit);
            if ((e = hipGetLastError()) != hipSuccess) return e;
            if (!same) {
                hipLaunchKernelGGL(planes_fn, dim3((unsigned)(nb * rsplit)), dim3(256), lds,
                                   stream, b, nb, n, ldb, pb, wpad, tb, rsplit);
                if ((e = hipGetLastError()) != hipSuccess) return e;
            }
        }
        {
            const int64_t items = (n - 1) + (wpad - words_upto(n - 1));
            hipLaunchKernelGGL(kendall_pastend_kernel, dim3(grid_for(items, 256)), dim3(256), 0,
                               stream, n, wpad, pe);
            if ((e = hipGetLastError()) != hipSuccess) return e;
        }
        KendallArgs args;
        args.pe = pe;
        args.pa = pa;
        args.pb = pb;
        args.na = na;
        args.nb = nb;
        args.wpad = wpad;
        args.count = count;
        args.ldc = nb;
        args.ties_a = ta;
        args.ties_b = tb;
        args.tiles_a = (na + KBM - 1) / KBM;
        args.tiles_b = (nb + KBM - 1) / KBM;
        args.ntiles = same ? args.tiles_a * (args.tiles_a + 1) / 2 : args.tiles_a * args.tiles_b;
        // split the words so that about 2048 workgroups run (8 per CU)
        int64_t splits = (2048 + args.ntiles - 1) / args.ntiles;
        const int64_t chunks = wpad / KKW;
        if (splits > chunks) splits = chunks;
        if (splits < 1) splits = 1;
        args.split_w = (chunks + splits - 1) / splits * KKW;
        // a workgroup's counts are 32-bit: fewer than 2^32 sample pairs
        // (32 per word) per word range
        const int64_t max_split_w = ((int64_t(1) << 27) - KKW) / KKW * KKW;
        if (args.split_w > max_split_w) args.split_w = max_split_w;
        splits = (wpad + args.split_w - 1) / args.split_w;
        const int64_t ids = args.ntiles * splits;
        const int64_t max_ids = int64_t(1) << 23;
        for (int64_t off = 0; off < ids; off += max_ids) {
            args.id_offset = off;
            args.ids_this = (ids - off) < max_ids ? (ids - off) : max_ids;
            if (same)
                hipLaunchKernelGGL(kendall_count_kernel<true>, dim3((unsigned)args.ids_this),
                                   dim3(KTHREADS), 0, stream, args);
            else
                hipLaunchKernelGGL(kendall_count_kernel<false>, dim3((unsigned)args.ids_this),
                                   dim3(KTHREADS), 0, stream, args);
            if ((e = hipGetLastError()) != hipSuccess) return e;
        }
    }
    // n < 2 leaves count null: no sample pair, tau = 0 / 0
    hipLaunchKernelGGL(kendall_finish_kernel, dim3(grid_for(na * nb, 256)), dim3(256), 0, stream,
                       count, nb, na, nb, n, same ? 1 : 0, tau, ldt);
    return hipGetLastError();
}

}  // namespace gt
