// gt_gram.hip -- split-precision MFMA Gram form of the euclidean distance.
//
// d(x_i, x_j)^2 = ||x_i||^2 + ||x_j||^2 - 2 x_i.x_j.  The dot products of a
// 256 x 256 tile of pairs run on the gfx950 matrix cores
// (v_mfma_f32_32x32x16_{bf16,f16}) with FP32 accumulation.  A pre-pass splits
// every FP32 input value once into 16-bit terms (stored in the workspace), and
// the tile kernel accumulates the products whose error terms matter:
//   BF16 (8-bit significand): three terms x = h + m + l, exact (x - h and
//     x - h - m are exact in FP32 and the last remainder has at most 8
//     significant bits); six products hh' + hm' + mh' + hl' + lh' + mm'; the
//     dropped ml' + lm' + ll' are below 2^-23 |x_c y_c| per feature.
//   FP16 (11-bit significand): two terms x = h + l (+ r, |r| <= 2^-22|x|),
//     three products hh' + hl' + lh'; dropped terms below 3 * 2^-22 |x_c y_c|.
//     Rows are first scaled by a power of two so that max |x| lies in
//     [2^14, 2^15) (keeps l out of the subnormal range and h below the FP16
//     maximum); the scale is undone exactly on the dot product.
// (A two-term BF16 split leaves ~3 * 2^-16 |x_c y_c| per feature, which no
// cancellation threshold can bring under 1e-5 for small k.)
//
// The Gram form loses relative accuracy when d^2 is small against
// ||x||^2 + ||y||^2 (cancellation).  Every pair with
//   d^2 < tau (||x||^2 + ||y||^2)
// is flagged in the epilogue and recomputed in the same workgroup from exact
// FP32 differences, sum (x_c - y_c)^2, with the FP32 rows re-streamed through
// LDS in 32-feature chunks (two-level summation).  Flagged pairs are processed
// through an LDS list of 1024 entries per pass, as many passes as needed, so
// any number of near-coincident pairs is handled.  tau is derived on the host
// from k (gram_tau) so that the Gram-form error stays below 1.6e-5 d^2.
//
// The reference computes this distance with exact differences only
// (inst/cuda/distance.cu:40-78); this path produces the same distance within
// the stated tolerance (max relative error <= 1e-5 against FP64).
#include "gt_common.h"
#include "gt_internal.h"
#include "gt_mfma.h"

namespace gt {
namespace {

constexpr int GBM = 256;        // tile side (points)
constexpr int GBK = 16;         // features per main-loop chunk (one MFMA k-step)
constexpr int KPAD = 32;        // the split terms are zero-padded to a multiple of this
constexpr int GROWB = GBK * 2;  // bytes of one row of one term in one chunk (two 16 B halves)
constexpr int GSUPER = 4;       // super-tile side in tiles (L2 locality)
constexpr int FBK = 32;         // features per fix-up chunk
constexpr int FLD = FBK + 4;    // fix-up LDS row length (floats, 16 B aligned rows)
constexpr int FCAP = 1024;      // flagged pairs per fix-up pass
// Gram-form values are finished with the hardware square root (v_sqrt_f32)
// plus one residual correction (sqrt_rn), which rounds correctly for normal
// inputs; d^2 below 2^-100 goes to the exact pass instead.
constexpr float GRAM_SQRT_MIN = 0x1p-100f;
constexpr int TERM_BYTES = GBM * GROWB;        // one split term of one side, one chunk (8 KB)
constexpr int FIX_BYTES = 2 * GBM * FLD * 4;   // fix-up staging, both sides
constexpr int LIST_BYTES = FCAP * 4;
constexpr int SIDE_BYTES = 4 * GBM * 4 + 16 * 4;  // norms A, B, inverse scales A, B, scan

// Wave layout of a 256 x 256 tile.  NW waves in a 4 (A rows) x NW/4 (B rows)
// grid; each wave holds 64 x (256 / (NW/4)) pairs as 2 x NB blocks of 32 x 32
// accumulators (NB = 8 / (NW/4)).  NW = 16: 64 x 64 per wave, 1024 threads,
// 128 VGPRs per lane.  NW = 8: 64 x 128 per wave, 512 threads, up to 256
// VGPRs per lane (the accumulators take 128).
template <int NW>
struct Waves {
    static constexpr int THREADS = NW * 64;
    static constexpr int WC = NW / 4;             // B-row groups
    static constexpr int NB = 8 / WC;             // 32-row accumulator blocks along B per wave
    static constexpr int FUNITS = 2 * GBM * FBK / 4 / THREADS;  // fix-up float4 units per thread
    static constexpr int FPER = FCAP / THREADS;   // fix-up list entries per thread
    static constexpr int GPW = 16 / NW;           // 32-row load groups per wave per term
};

// LDS: [split terms, three stages | (aliased) fix-up staging + pair list]
// [side data]
template <int NT>
struct TileLds {
    // main-loop ring depth (two stages for BF16 measured no faster: 15.67 vs
    // 15.55 ms on the 65536 x 512 mixture)
    static constexpr int NSTAGE = 3;
    static constexpr int STAGE = 2 * NT * TERM_BYTES;
    static constexpr int MAIN = NSTAGE * STAGE;
    static constexpr int FIX = FIX_BYTES + LIST_BYTES;
    static constexpr int REGION = MAIN > FIX ? MAIN : FIX;
    static constexpr int SIDE = REGION;
    static constexpr int BYTES = REGION + SIDE_BYTES;
};

struct GramArgs {
    const float *x;        // FP32 points (exact passes)
    int64_t n, k, ldx;
    const unsigned short *terms;  // NT x n x kp split terms
    int64_t kp;            // k padded to KPAD
    const float *norms;    // ||x_i - c||^2 (FP32)
    const float *iscale;   // FP16 path: inverse power-of-two row scale
    float *out;
    int64_t ldd;
    int64_t T;             // tiles per side
    int64_t id_offset, ids_this;
    float tau;
};

// Pre-pass: split every centred value x - c once.  Term t of row r, feature
// c, at terms[(t * n + r) * kp + c]; features k..kp-1 are zero.  One thread
// per 8 consecutive features (one 16 B store per term).
template <int SPLIT, bool VEC4>
__global__ __launch_bounds__(256) void split_kernel(const float *__restrict__ x, int64_t n,
                                                    int64_t k, int64_t ldx, int64_t kp,
                                                    const float *__restrict__ center,
                                                    const float *__restrict__ scales,
                                                    unsigned short *__restrict__ terms)
{
    constexpr int NT = Split<SPLIT>::NT;
    const int64_t per_row = kp / 8;
    const int64_t units = n * per_row;
    for (int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; u < units;
         u += (int64_t)gridDim.x * blockDim.x) {
        int64_t r, c0;
        if (units <= 0xffffffffll) {
            // 32-bit division (the 64-bit one is a long instruction sequence)
            const uint32_t uu = (uint32_t)u, pr = (uint32_t)per_row;
            r = uu / pr;
            c0 = (int64_t)(uu - (uint32_t)r * pr) * 8;
        } else {
            r = u / per_row;
            c0 = (u % per_row) * 8;
        }
        const float sc = (SPLIT == SPLIT_FP16) ? scales[r] : 1.0f;
        float v[8];
        if (VEC4 && c0 + 8 <= k) {
            const float4 a = *reinterpret_cast<const float4 *>(x + r * ldx + c0);
            const float4 b = *reinterpret_cast<const float4 *>(x + r * ldx + c0 + 4);
            const float4 ca = *reinterpret_cast<const float4 *>(center + c0);
            const float4 cb = *reinterpret_cast<const float4 *>(center + c0 + 4);
            v[0] = a.x - ca.x; v[1] = a.y - ca.y; v[2] = a.z - ca.z; v[3] = a.w - ca.w;
            v[4] = b.x - cb.x; v[5] = b.y - cb.y; v[6] = b.z - cb.z; v[7] = b.w - cb.w;
        } else {
#pragma unroll
            for (int c = 0; c < 8; ++c) v[c] = (c0 + c < k) ? x[r * ldx + c0 + c] - center[c0 + c] : 0.0f;
        }
        unsigned short t[8][NT];
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            if (SPLIT == SPLIT_FP16) v[c] *= sc;
            Split<SPLIT>::split(v[c], t[c]);
        }
#pragma unroll
        for (int q = 0; q < NT; ++q) {
            uint4 w;
            w.x = t[0][q] | ((unsigned)t[1][q] << 16);
            w.y = t[2][q] | ((unsigned)t[3][q] << 16);
            w.z = t[4][q] | ((unsigned)t[5][q] << 16);
            w.w = t[6][q] | ((unsigned)t[7][q] << 16);
            *reinterpret_cast<uint4 *>(terms + ((int64_t)q * n + r) * kp + c0) = w;
        }
    }
}

// Exclusive scan of one int per thread over a workgroup of NW waves.
template <int NW>
__device__ __forceinline__ int block_exclusive_scan(int v, int *tmp, int &total)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int incl = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int t = __shfl_up(incl, d, 64);
        if (lane >= d) incl += t;
    }
    if (lane == 63) tmp[wave] = incl;
    __syncthreads();
    int before = 0;
    total = 0;
#pragma unroll
    for (int w = 0; w < NW; ++w) {
        const int c = tmp[w];
        if (w < wave) before += c;
        total += c;
    }
    __syncthreads();
    return before + incl - v;
}

// Wait until at most N vector memory operations of this wave are outstanding.
template <int N>
__device__ __forceinline__ void wait_vmcnt()
{
    asm volatile("s_waitcnt vmcnt(%0)" ::"n"(N) : "memory");
}

typedef __attribute__((address_space(3))) void lds_void_t;
typedef const __attribute__((address_space(1))) void gbl_void_t;

// Main-loop LDS image of one chunk: side A (rows j) terms 0..NT-1, then side
// B (rows i), each 256 rows x 32 B; the two 16 B halves of row R are swapped
// when swz(R) = ((R >> 2) ^ (R >> 3)) & 1, so that the b128 operand reads of
// 8 or 16 consecutive rows hit distinct banks without padding.
// global_load_lds writes lane l's 16 B at the wave's LDS base + 16 l, so each
// wave instruction fills 32 rows of one term.  The 16 groups of 32 rows
// (8 per side) go to the waves round robin: wave w fills groups w, w + NW,
// ... (group g: side g / 8, rows 32 (g % 8) ..).  Rows past n are clamped to
// row n - 1: they only feed pairs that are never stored.
template <int NW>
__device__ __forceinline__ void chunk_sources(const GramArgs &args, int64_t i0, int64_t j0,
                                              int wave, int lane,
                                              const unsigned short *(&gsrc)[Waves<NW>::GPW])
{
    const int rr = lane >> 1;
    const int half = (lane & 1) ^ (((rr >> 2) ^ (rr >> 3)) & 1);
#pragma unroll
    for (int q = 0; q < Waves<NW>::GPW; ++q) {
        const int g = wave + q * NW;
        const int64_t r = ((g >> 3) ? i0 : j0) + (g & 7) * 32 + rr;
        gsrc[q] = args.terms + (r < args.n ? r : args.n - 1) * args.kp + half * 8;
    }
}

template <int NT, int NW>
__device__ __forceinline__ void issue_chunk(const GramArgs &args,
                                            const unsigned short *const (&gsrc)[Waves<NW>::GPW],
                                            unsigned char *stage, int wave, int chunk)
{
    const int64_t term_stride = args.n * args.kp;
    const int64_t kb = (int64_t)chunk * GBK;
#pragma unroll
    for (int q = 0; q < Waves<NW>::GPW; ++q) {
        const int g = wave + q * NW;
        unsigned char *dst = stage + (g >> 3) * NT * TERM_BYTES + (g & 7) * 1024;
#pragma unroll
        for (int t = 0; t < NT; ++t)
            __builtin_amdgcn_global_load_lds((gbl_void_t *)(gsrc[q] + t * term_stride + kb),
                                             (lds_void_t *)(dst + t * TERM_BYTES), 16, 0, 0);
    }
}

// One chunk of MFMA products for this wave's pairs (acc[a][b]: A rows
// wr*64 + a*32.., B rows wc*NB*32 + b*32..); sg = stage + the lane's operand
// offset.
template <int SPLIT, int NB>
__device__ __forceinline__ void mfma_chunk(f32x16_t (&acc)[2][NB], const unsigned char *sg, int wr,
                                           int wc)
{
    typedef typename Split<SPLIT>::vec vec;
    constexpr int NT = Split<SPLIT>::NT;
    vec av[NT][2];
#pragma unroll
    for (int t = 0; t < NT; ++t)
#pragma unroll
        for (int a = 0; a < 2; ++a)
            av[t][a] = *reinterpret_cast<const vec *>(sg + t * TERM_BYTES + (wr * 64 + a * 32) * GROWB);
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        vec bv[NT];
#pragma unroll
        for (int t = 0; t < NT; ++t)
            bv[t] = *reinterpret_cast<const vec *>(sg + (NT + t) * TERM_BYTES +
                                                   (wc * NB * 32 + b * 32) * GROWB);
        // hh' first, then the cross terms in decreasing magnitude
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

// Norms (and FP16 inverse row scales) of the tile's rows into LDS.
template <int SPLIT>
__device__ __forceinline__ void load_side(const GramArgs &args, int64_t i0, int64_t j0, float *nA,
                                          float *nB, float *iA, float *iB)
{
    for (int t = threadIdx.x; t < 2 * GBM; t += blockDim.x) {
        if (t < GBM) {
            const int64_t r = j0 + t;
            nA[t] = r < args.n ? args.norms[r] : 0.0f;
            if (SPLIT == SPLIT_FP16) iA[t] = r < args.n ? args.iscale[r] : 1.0f;
        } else {
            const int64_t r = i0 + (t - GBM);
            nB[t - GBM] = r < args.n ? args.norms[r] : 0.0f;
            if (SPLIT == SPLIT_FP16) iB[t - GBM] = r < args.n ? args.iscale[r] : 1.0f;
        }
    }
}

// Epilogue of one tile: d^2 = ||x||^2 + ||y||^2 - 2g for this wave's pairs;
// every valid pair that is not flagged (d^2 >= tau (||x||^2 + ||y||^2)) is
// stored.  Flag bits: index (a * NB + b) * 16 + r marks accumulator element
// r of block (a, b), in fm[index / 64].  32x32 accumulator layout: element r
// of lane l is row 8(r/4) + 4(l/32) + r%4 (A side, j) and column l%32 (B
// side, i); lanes 0..31 hold 32 consecutive i of one j, so packed stores are
// contiguous runs.
template <int SPLIT, int LAYOUT, int NB>
__device__ __forceinline__ void gram_epilogue(const f32x16_t (&acc)[2][NB], const GramArgs &args,
                                              int64_t i0, int64_t j0, bool interior, int wr,
                                              int wc, int lane, const float *nA, const float *nB,
                                              const float *iA, const float *iB,
                                              unsigned long long (&fm)[NB / 2])
{
    const int64_t n = args.n;
    const int li0 = wc * NB * 32 + (lane & 31);
    float ni[NB], sb[NB];
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        ni[b] = nB[li0 + b * 32];
        sb[b] = (SPLIT == SPLIT_FP16) ? iB[li0 + b * 32] : 1.0f;
    }
#pragma unroll
    for (int q = 0; q < NB / 2; ++q) fm[q] = 0ull;
    const float tau = args.tau;
    if (interior) {
        // every pair valid, i > j: stored branch-free.  Rows jw + d (d < 64)
        // of this wave: element offsets from the start of row jw fit in 32
        // bits (the host limits n to below 2^24), and the row-dependent part
        // of each offset is wave-uniform
        const int64_t jw = j0 + wr * 64;
        float *const outw =
            args.out + ((LAYOUT == OUT_PACKED) ? packed_base(n, jw) : jw * args.ldd) + i0;
        const unsigned cw = (LAYOUT == OUT_PACKED) ? (unsigned)(n - 1 - jw) : (unsigned)args.ldd;
        const bool hi = (lane >> 5) != 0;
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const unsigned dc = a * 32 + 8 * (r >> 2) + (r & 3);
                const unsigned base_u = (LAYOUT == OUT_PACKED) ? dc * cw - dc * (dc + 1) / 2 : dc * cw;
                const unsigned step_u = (LAYOUT == OUT_PACKED) ? 4 * cw - 4 * dc - 10 : 4 * cw;
                const unsigned ro = base_u + (hi ? step_u : 0u) + (unsigned)li0;
                const int lj = wr * 64 + (int)dc + (hi ? 4 : 0);
                const float nj = nA[lj];
                const float sj = (SPLIT == SPLIT_FP16) ? iA[lj] : 1.0f;
#pragma unroll
                for (int b = 0; b < NB; ++b) {
                    float g = acc[a][b][r];
                    if (SPLIT == SPLIT_FP16) g = (g * sj) * sb[b];
                    const float ss = nj + ni[b];
                    const float d2 = fmaf(-2.0f, g, ss);
                    const bool ok = d2 >= fmaf(tau, ss, GRAM_SQRT_MIN);
                    const int bit = (a * NB + b) * 16 + r;
                    fm[bit >> 6] |= ok ? 0ull : (1ull << (bit & 63));
                    const float v = sqrt_rn(d2);
                    outw[ro + (unsigned)(b * 32)] = v;
                    if (LAYOUT == OUT_FULL) args.out[(i0 + li0 + b * 32) * args.ldd + jw + lj - wr * 64] = v;
                }
            }
    } else {
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int lj = wr * 64 + a * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
                const int64_t j = j0 + lj;
                const float nj = nA[lj];
                const float sj = (SPLIT == SPLIT_FP16) ? iA[lj] : 1.0f;
                float *rowp = args.out + ((LAYOUT == OUT_PACKED) ? packed_base(n, j) : j * args.ldd) +
                              i0 + li0;
#pragma unroll
                for (int b = 0; b < NB; ++b) {
                    float g = acc[a][b][r];
                    // undo the row scales one at a time (their product can overflow)
                    if (SPLIT == SPLIT_FP16) g = (g * sj) * sb[b];
                    const float ss = nj + ni[b];
                    const float d2 = fmaf(-2.0f, g, ss);
                    const bool flag = !(d2 >= fmaxf(tau * ss, GRAM_SQRT_MIN));
                    const int li = li0 + b * 32;
                    const bool valid = i0 + li < n && i0 + li > j;
                    if (valid) {
                        if (!flag) {
                            const float v = sqrt_rn(fmaxf(d2, 0.0f));
                            rowp[b * 32] = v;
                            if (LAYOUT == OUT_FULL) args.out[(i0 + li) * args.ldd + j] = v;
                        } else {
                            const int bit = (a * NB + b) * 16 + r;
                            fm[bit >> 6] |= 1ull << (bit & 63);
                        }
                    }
                    if (LAYOUT == OUT_FULL && i0 + li == j && j < n) rowp[b * 32] = 0.0f;
                }
            }
    }
}

// Tile position of flag index `bit` of this lane: (A row << 16) | B row.
template <int NB>
__device__ __forceinline__ unsigned flag_entry(int bit, int wr, int li0, int lane)
{
    const int a = bit / (NB * 16), b = (bit / 16) % NB, r = bit & 15;
    const int lj = wr * 64 + a * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
    return ((unsigned)lj << 16) | (unsigned)(li0 + b * 32);
}

// Exact distance of a flagged pair (entry as flag_entry) of tile (i0, j0).
template <int LAYOUT>
__device__ __forceinline__ void store_exact(const GramArgs &args, int64_t i0, int64_t j0,
                                            unsigned ent, float sum)
{
    const int64_t j = j0 + (ent >> 16), i = i0 + (ent & 0xffffu);
    const float v = sqrtf(sum);
    if (LAYOUT == OUT_PACKED) {
        args.out[packed_base(args.n, j) + i] = v;
    } else {
        args.out[j * args.ldd + i] = v;
        args.out[i * args.ldd + j] = v;
    }
}

// Exact FP32 recompute of the tile's flagged pairs: the FP32 rows of both
// sides are re-streamed through LDS in 32-feature chunks (two-level sums);
// the pairs go through an LDS list of FCAP entries per pass.  myoff: this
// lane's rank among the flagged pairs; total: their number.  The caller has
// let the epilogue's stores land.
template <int LAYOUT, bool VEC4, int NW>
__device__ __forceinline__ void gram_fixup(const GramArgs &args, unsigned char *smem, int64_t i0,
                                           int64_t j0,
                                           const unsigned long long (&fm)[Waves<NW>::NB / 2],
                                           int myoff, int total, int wr, int li0, int lane)
{
    typedef Waves<NW> W;
    constexpr int THREADS = W::THREADS, FUNITS = W::FUNITS, FPER = W::FPER, NB = W::NB;
    const int tid = threadIdx.x;
    float *fS = reinterpret_cast<float *>(smem);
    unsigned int *list = reinterpret_cast<unsigned int *>(smem + FIX_BYTES);
    const int64_t n = args.n, k = args.k, ldx = args.ldx;
    // Staging unit e = tid + THREADS q (q < FUNITS): tile row e >> 3 (A side
    // rows first, then B side), features (e & 7) * 4..+3 of a chunk; the
    // unit's column is (tid & 7) * 4 for every q.  The accumulators are dead
    // here, so their registers hold a ring of FRING chunks in flight (the
    // pass is bound by load latency, not bandwidth).
    constexpr int FRING = 2;
    int frow[FUNITS];
#pragma unroll
    for (int q = 0; q < FUNITS; ++q) {
        const int row = (tid + THREADS * q) >> 3;
        const int64_t g = row < GBM ? j0 + row : i0 + (row - GBM);
        frow[q] = (int)(g < n ? g : n - 1);
    }
    const float *const fx = args.x + (tid & 7) * 4;
    const int nf = (int)((k + FBK - 1) / FBK);
    float4 fr0[FUNITS], fr1[FUNITS];
#define GT_FLOAD(buf, chunk)                                                               \
    do {                                                                                   \
        const int64_t fb_ = (int64_t)(chunk) * FBK;                                        \
        const int64_t kk = fb_ + (tid & 7) * 4;                                            \
        _Pragma("unroll") for (int q = 0; q < FUNITS; ++q) {                               \
            const float *src_ = fx + (int64_t)frow[q] * ldx + fb_;                         \
            if (VEC4) {                                                                    \
                buf[q] = kk < k ? *reinterpret_cast<const float4 *>(src_)                  \
                                : make_float4(0.f, 0.f, 0.f, 0.f);                         \
            } else {                                                                       \
                float v_[4];                                                               \
                _Pragma("unroll") for (int c_ = 0; c_ < 4; ++c_) v_[c_] =                  \
                    (kk + c_ < k) ? src_[c_] : 0.0f;                                       \
                buf[q] = make_float4(v_[0], v_[1], v_[2], v_[3]);                          \
            }                                                                              \
        }                                                                                  \
    } while (0)
    // one chunk: registers -> LDS, refill the registers FRING chunks ahead,
    // then accumulate this thread's flagged pairs over the chunk
#define GT_FSTEP(buf, chunk)                                                               \
    do {                                                                                   \
        __syncthreads();                                                                   \
        _Pragma("unroll") for (int q = 0; q < FUNITS; ++q) {                               \
            const int e = tid + THREADS * q;                                               \
            *reinterpret_cast<float4 *>(fS + (e >> 3) * FLD + (e & 7) * 4) = buf[q];       \
        }                                                                                  \
        __syncthreads();                                                                   \
        if ((chunk) + FRING < nf) GT_FLOAD(buf, (chunk) + FRING);                          \
        _Pragma("unroll") for (int c = 0; c < FPER; ++c) {                                 \
            if (tid + THREADS * c < cnt) {                                                 \
                const float4 *ra = reinterpret_cast<const float4 *>(fS + (ent[c] >> 16) * FLD); \
                const float4 *rb =                                                         \
                    reinterpret_cast<const float4 *>(fS + (GBM + (ent[c] & 0xffffu)) * FLD); \
                float part = 0.0f;                                                         \
                _Pragma("unroll") for (int q = 0; q < FBK / 4; ++q) {                      \
                    const float4 u = ra[q], v = rb[q];                                     \
                    float t = u.x - v.x;                                                   \
                    part = fmaf(t, t, part);                                               \
                    t = u.y - v.y;                                                         \
                    part = fmaf(t, t, part);                                               \
                    t = u.z - v.z;                                                         \
                    part = fmaf(t, t, part);                                               \
                    t = u.w - v.w;                                                         \
                    part = fmaf(t, t, part);                                               \
                }                                                                          \
                sum[c] += part;                                                            \
            }                                                                              \
        }                                                                                  \
    } while (0)
    for (int base = 0; base < total; base += FCAP) {
        {
            int rank = myoff;
#pragma unroll
            for (int qw = 0; qw < NB / 2; ++qw) {
                unsigned long long m = fm[qw];
                while (m) {
                    const int bit = __builtin_ctzll(m);
                    m &= m - 1;
                    if (rank >= base && rank < base + FCAP)
                        list[rank - base] = flag_entry<NB>(qw * 64 + bit, wr, li0, lane);
                    ++rank;
                }
            }
        }
        const int cnt = (total - base) < FCAP ? (total - base) : FCAP;
        float sum[FPER];
        unsigned ent[FPER];
#pragma unroll
        for (int c = 0; c < FPER; ++c) sum[c] = 0.0f;
        GT_FLOAD(fr0, 0);
        if (nf > 1) GT_FLOAD(fr1, 1);
        __syncthreads();
#pragma unroll
        for (int c = 0; c < FPER; ++c) {
            const int e = tid + THREADS * c;
            ent[c] = e < cnt ? list[e] : 0u;
        }
        for (int ch = 0; ch < nf; ch += FRING) {
            GT_FSTEP(fr0, ch);
            if (ch + 1 < nf) GT_FSTEP(fr1, ch + 1);
        }
#pragma unroll
        for (int c = 0; c < FPER; ++c)
            if (tid + THREADS * c < cnt) store_exact<LAYOUT>(args, i0, j0, ent[c], sum[c]);
        __syncthreads();
    }
#undef GT_FSTEP
#undef GT_FLOAD
}

// One workgroup per 256 x 256 tile: main loop, epilogue, then the exact
// pass of the flagged pairs.
template <int SPLIT, int LAYOUT, bool VEC4, int NW>
__global__ __launch_bounds__(NW * 64) void gram_tile_kernel(GramArgs args)
{
    typedef Waves<NW> W;
    constexpr int NT = Split<SPLIT>::NT, NB = W::NB, GPW = W::GPW;
    typedef TileLds<NT> L;
    __shared__ __attribute__((aligned(1024))) unsigned char smem[L::BYTES];
    float *nA = reinterpret_cast<float *>(smem + L::SIDE);
    float *nB = nA + GBM, *iA = nB + GBM, *iB = iA + GBM;
    int *scan_tmp = reinterpret_cast<int *>(iB + GBM);

    const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wr = wave / W::WC, wc = wave % W::WC;

    // Super-tiled lower-triangular schedule: super-tile s covers GSUPER x
    // GSUPER tiles; tiles above the diagonal or past the edge exit at once.
    const int64_t id = args.id_offset + xcd_remap(blockIdx.x, args.ids_this);
    const int64_t st = id / (GSUPER * GSUPER);
    const int w = (int)(id % (GSUPER * GSUPER));
    int SI, SJ;
    tri_tile(st, SI, SJ);
    const int64_t I = (int64_t)SI * GSUPER + w / GSUPER;
    const int64_t J = (int64_t)SJ * GSUPER + w % GSUPER;
    if (I >= args.T || J > I) return;
    const int64_t i0 = I * GBM, j0 = J * GBM;
    load_side<SPLIT>(args, i0, j0, nA, nB, iA, iB);

    const unsigned short *gsrc[GPW];
    chunk_sources<NW>(args, i0, j0, wave, lane, gsrc);
    const int lrow = lane & 31;
    const int roff = lrow * GROWB + (((lane >> 5) ^ (((lrow >> 2) ^ (lrow >> 3)) & 1)) * 16);
    f32x16_t acc[2][NB];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < NB; ++b)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[a][b][r] = 0.0f;

    const int nchunks = (int)(args.kp / GBK);
    // NSTAGE stages: chunks c + 1 .. c + NSTAGE - 1 in flight while c is
    // computed
    constexpr int NS = L::NSTAGE;
#pragma unroll
    for (int c = 0; c < NS - 1; ++c)
        if (c < nchunks) issue_chunk<NT, NW>(args, gsrc, smem + c * L::STAGE, wave, c);
    for (int c = 0; c < nchunks; ++c) {
        // chunk c landed when at most the next chunk's loads of this wave
        // are outstanding (vector memory loads complete in order)
        if (NS == 3 && c + 1 < nchunks)
            wait_vmcnt<NT * GPW>();
        else
            wait_vmcnt<0>();
        // every wave's part of chunk c is in LDS, and every wave finished
        // chunk c-1, whose stage the next issue overwrites.  A bare s_barrier:
        // __syncthreads() would also drain the later chunks' loads.
        __builtin_amdgcn_s_barrier();
        asm volatile("" ::: "memory");
        if (c + NS - 1 < nchunks)
            issue_chunk<NT, NW>(args, gsrc, smem + ((c + NS - 1) % NS) * L::STAGE, wave, c + NS - 1);
        mfma_chunk<SPLIT, NB>(acc, smem + (c % NS) * L::STAGE + roff, wr, wc);
    }
    // the fix-up below reuses the stage memory
    __syncthreads();

    const bool interior = (I > J) && (i0 + GBM <= args.n);
    unsigned long long fm[NB / 2];
    gram_epilogue<SPLIT, LAYOUT, NB>(acc, args, i0, j0, interior, wr, wc, lane, nA, nB, iA, iB, fm);
    int cnt = 0;
#pragma unroll
    for (int q = 0; q < NB / 2; ++q) cnt += __popcll(fm[q]);
    int total;
    const int myoff = block_exclusive_scan<NW>(cnt, scan_tmp, total);
    if (total == 0) return;
    // the epilogue's stores of flagged pairs land before the exact values
    wait_vmcnt<0>();
    __syncthreads();
    gram_fixup<LAYOUT, VEC4, NW>(args, smem, i0, j0, fm, myoff, total, wr,
                                 wc * NB * 32 + (lane & 31), lane);
}

// Column centre c_j of the Gram form.  Euclidean distances do not change
// when every point is moved by the same vector, but the Gram form's error
// scales with ||x||^2 + ||y||^2: centring the columns makes that as small as
// the data's spread allows (U[0,1) data: 4x; N(0,1) + 100: 10^4 x).  c_j is
// the column mean rounded to a power-of-two grid no coarser than 1/64 of the
// column's range, so that x - c_j is often exact (integer columns keep
// integer centres); a constant column takes its value; a column holding a
// NaN or an infinity is not moved.
constexpr int CS_ROWS = 256;    // rows per partial-statistics block

__global__ __launch_bounds__(256) void colstats_partial_kernel(const float *__restrict__ x,
                                                               int64_t n, int64_t k, int64_t ldx,
                                                               float *__restrict__ psum,
                                                               float *__restrict__ pmin,
                                                               float *__restrict__ pmax)
{
    const int64_t r0 = (int64_t)blockIdx.x * CS_ROWS;
    const int64_t r1 = (r0 + CS_ROWS < n) ? r0 + CS_ROWS : n;
    for (int64_t c = (int64_t)blockIdx.y * 256 + threadIdx.x; c < k; c += (int64_t)gridDim.y * 256) {
        float sum = 0.0f, mn = __builtin_inff(), mx = -__builtin_inff();
#pragma unroll 8
        for (int64_t r = r0; r < r1; ++r) {
            const float v = x[r * ldx + c];
            sum += v;              // a NaN or an infinity makes the sum non-finite
            mn = fminf(mn, v);
            mx = fmaxf(mx, v);
        }
        psum[blockIdx.x * k + c] = sum;
        pmin[blockIdx.x * k + c] = mn;
        pmax[blockIdx.x * k + c] = mx;
    }
}

// Column statistics from the partials: 16 columns per workgroup, 16
// threads per column, each over every 16th partial block; the 16 partial
// results are combined in a fixed order.
constexpr int CS_COLS = 16, CS_SLICES = 16;

__global__ __launch_bounds__(256) void colstats_finish_kernel(const float *__restrict__ psum,
                                                              const float *__restrict__ pmin,
                                                              const float *__restrict__ pmax,
                                                              int64_t blocks, int64_t n, int64_t k,
                                                              float *__restrict__ center)
{
    __shared__ float red[3][CS_SLICES][CS_COLS];
    const int cl = threadIdx.x % CS_COLS, sl = threadIdx.x / CS_COLS;
    for (int64_t cb = (int64_t)blockIdx.x * CS_COLS; cb < k; cb += (int64_t)gridDim.x * CS_COLS) {
        const int64_t c = cb + cl;
        float sum = 0.0f, mn = __builtin_inff(), mx = -__builtin_inff();
        if (c < k) {
            for (int64_t b = sl; b < blocks; b += CS_SLICES) {
                sum += psum[b * k + c];
                mn = fminf(mn, pmin[b * k + c]);
                mx = fmaxf(mx, pmax[b * k + c]);
            }
        }
        red[0][sl][cl] = sum;
        red[1][sl][cl] = mn;
        red[2][sl][cl] = mx;
        __syncthreads();
        if (sl == 0 && c < k) {
            sum = 0.0f;
            for (int q = 0; q < CS_SLICES; ++q) {
                sum += red[0][q][cl];
                mn = fminf(mn, red[1][q][cl]);
                mx = fmaxf(mx, red[2][q][cl]);
            }
            float cc = 0.0f;
            const float range = mx - mn;
            if (isfinite(sum) && isfinite(range)) {
                if (range == 0.0f) {
                    cc = mn;
                } else {
                    const float mean = sum / (float)n;
                    const float g = ldexpf(1.0f, ilogbf(range) - 6);   // <= range / 64
                    const float q = mean / g;
                    cc = (fabsf(q) < 0x1p24f) ? rintf(q) * g : mean;
                }
            }
            center[c] = cc;
        }
        __syncthreads();
    }
}

// ||x_i - c||^2 per row (and, for FP16, the power-of-two row scale and its
// inverse): one wave per row.  The squares are summed with compensation
// (each product split exactly by an fma, each addition's rounding error
// recovered by TwoSum), so the norm is within about 2u of exact whatever k.
__global__ __launch_bounds__(256) void row_prep_kernel(const float *__restrict__ x, int64_t n,
                                                       int64_t k, int64_t ldx,
                                                       const float *__restrict__ center,
                                                       float *__restrict__ norms,
                                                       float *__restrict__ scales,
                                                       float *__restrict__ iscales)
{
#pragma clang fp contract(off)
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int64_t row = (int64_t)blockIdx.x * 4 + wave; row < n; row += (int64_t)gridDim.x * 4) {
        const float *xr = x + row * ldx;
        float s = 0.0f, comp = 0.0f, m = 0.0f;
        for (int64_t c = lane; c < k; c += 64) {
            const float v = xr[c] - center[c];
            const float p = v * v;
            const float pe = fmaf(v, v, -p);          // v * v = p + pe exactly
            const float t = s + p;
            const float bp = t - s;
            comp += ((s - (t - bp)) + (p - bp)) + pe;  // s + p = t + (TwoSum error)
            s = t;
            m = fmaxf(m, fabsf(v));
        }
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) {
            const float so = __shfl_xor(s, d, 64), co = __shfl_xor(comp, d, 64);
            const float t = s + so;
            const float bp = t - s;
            comp = (comp + co) + ((s - (t - bp)) + (so - bp));
            s = t;
            m = fmaxf(m, __shfl_xor(m, d, 64));
        }
        if (lane == 0) {
            norms[row] = s + comp;
            if (scales) {
                // power of two that brings max|x - c| into [2^14, 2^15)
                int e = 0;
                if (m > 0.0f) {
                    e = ilogbf(m);
                    if (e < -100) e = -100;
                    if (e > 100) e = 100;
                }
                scales[row] = ldexpf(1.0f, 14 - e);
                iscales[row] = ldexpf(1.0f, e - 14);
            }
        }
    }
}

// 16 waves (64 x 64 pairs each).  8 waves of 64 x 128 pairs (512 threads,
// about 200 VGPRs) give the same results bit for bit and measured slower:
// 16.8 vs 15.9 ms on the 65536 x 512 mixture, 12.8 vs 12.5 ms on N(0, 1)
// (BF16); the exact pass has half the loads in flight.
constexpr int kGramWaves = 16;

template <int SPLIT, int LAYOUT, bool VEC4>
hipError_t launch_tiles(GramArgs &args, int64_t ids, hipStream_t stream)
{
    const int64_t max_ids = int64_t(1) << 23;
    for (int64_t off = 0; off < ids; off += max_ids) {
        args.id_offset = off;
        args.ids_this = (ids - off) < max_ids ? (ids - off) : max_ids;
        hipLaunchKernelGGL((gram_tile_kernel<SPLIT, LAYOUT, VEC4, kGramWaves>),
                           dim3((unsigned)args.ids_this), dim3(kGramWaves * 64), 0, stream, args);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

template <int SPLIT>
hipError_t launch_split(int layout, bool vec4, GramArgs &args, const float *center,
                        const float *scales, int64_t ids, hipStream_t stream)
{
    {
        const int64_t units = args.n * (args.kp / 8);
        int64_t blocks = (units + 255) / 256;
        if (blocks > 65536) blocks = 65536;
        if (vec4)
            hipLaunchKernelGGL((split_kernel<SPLIT, true>), dim3((unsigned)blocks), dim3(256), 0, stream,
                               args.x, args.n, args.k, args.ldx, args.kp, center, scales,
                               const_cast<unsigned short *>(args.terms));
        else
            hipLaunchKernelGGL((split_kernel<SPLIT, false>), dim3((unsigned)blocks), dim3(256), 0,
                               stream, args.x, args.n, args.k, args.ldx, args.kp, center, scales,
                               const_cast<unsigned short *>(args.terms));
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    if (layout == OUT_PACKED)
        return vec4 ? launch_tiles<SPLIT, OUT_PACKED, true>(args, ids, stream)
                    : launch_tiles<SPLIT, OUT_PACKED, false>(args, ids, stream);
    return vec4 ? launch_tiles<SPLIT, OUT_FULL, true>(args, ids, stream)
                : launch_tiles<SPLIT, OUT_FULL, false>(args, ids, stream);
}

size_t align256(size_t v) { return (v + 255) / 256 * 256; }

int64_t padded_k(int64_t k) { return (k + KPAD - 1) / KPAD * KPAD; }

int split_terms(int precision) { return precision == 1 ? 3 : 2; }

}  // namespace

int64_t colstats_blocks(int64_t n) { return (n + CS_ROWS - 1) / CS_ROWS; }

// Workspace: split terms (NT x n x kp 16-bit values), then norms, scales and
// inverse scales (n floats each), the column centres (k floats) and the
// partial column statistics (3 x ceil(n / 256) x k floats).
size_t gram_workspace_bytes(int64_t n, int64_t k, int precision)
{
    const size_t terms = align256((size_t)split_terms(precision) * (size_t)n *
                                  (size_t)padded_k(k) * sizeof(unsigned short));
    return terms + 3 * align256((size_t)n * sizeof(float)) + align256((size_t)k * sizeof(float)) +
           3 * align256((size_t)colstats_blocks(n) * (size_t)k * sizeof(float));
}

// Cancellation threshold: pairs with d^2 < tau (||x||^2 + ||y||^2) (norms
// of the centred rows) are recomputed from exact differences of the original
// rows.  The Gram-form error of d^2, in units of ||x||^2 + ||y||^2
// (u = 2^-24):
//   split    BF16 three terms: 4u; FP16 two terms: 12u (= 3 * 2^-22)
//            (|dg| <= c sum|x_c y_c| <= c (||x||^2+||y||^2)/2 and |d(d^2)| = 2|dg|)
//   accumulation of the MFMA partial sums, one FP32 rounding per product
//            instruction: statistical bound 4 sqrt(#adds) u
//   norms    4u for the two compensated row norms together
//   final    3u for ||x||^2 + ||y||^2 - 2g
// Relative to d^2 of a pair that is not flagged (d^2 >= tau (...)), that is
// at most E / tau.  Centring moves each difference x_c - y_c by at most
// u (|x~_c| + |y~_c|), so d^2 by at most 2u d (||x~|| + ||y~||) <=
// 2u sqrt(2 / tau) d^2 <= 8u d^2 for tau >= 1/8.  tau is chosen so that
// E / tau + 8u = 1.6e-5, the allowed relative error of d^2 (8e-6 on d).
float gram_tau(int precision, int64_t k)
{
    const double u = 1.0 / 16777216.0;
    const bool bf16 = (precision == 1);
    const double split = bf16 ? 4.0 * u : 12.0 * u;
    const double nadds = (bf16 ? 6.0 : 3.0) * (double)((k + 15) / 16);
    const double accum = 4.0 * sqrt(nadds) * u;
    const double total = split + accum + 4.0 * u + 3.0 * u;
    const double tau = total / (1.6e-5 - 8.0 * u);
    return (float)(tau > 0.125 ? tau : 0.125);
}

// Above this tau nearly every pair would be recomputed: the request is
// computed by the FP32 difference kernel instead.
constexpr float GRAM_TAU_MAX = 0.9f;

hipError_t launch_gram(int precision, int layout, const float *x, int64_t n, int64_t k,
                       int64_t ldx, float *out, int64_t ldd, void *workspace, hipStream_t stream)
{
    if (n < 2 && layout == OUT_PACKED) return hipSuccess;
    const float tau = gram_tau(precision, k);
    if (tau > GRAM_TAU_MAX)
        return launch_plain(OP_EUCLID, layout, x, n, ldx, x, n, ldx, k, 2.0f, out, ldd, stream);
    const int64_t kp = padded_k(k);
    char *ws = static_cast<char *>(workspace);
    unsigned short *terms = reinterpret_cast<unsigned short *>(ws);
    ws += align256((size_t)split_terms(precision) * (size_t)n * (size_t)kp * sizeof(unsigned short));
    float *norms = reinterpret_cast<float *>(ws);
    ws += align256((size_t)n * sizeof(float));
    float *scales = reinterpret_cast<float *>(ws);
    ws += align256((size_t)n * sizeof(float));
    float *iscales = reinterpret_cast<float *>(ws);
    ws += align256((size_t)n * sizeof(float));
    float *center = reinterpret_cast<float *>(ws);
    ws += align256((size_t)k * sizeof(float));
    const int64_t csb = colstats_blocks(n);
    const size_t pbytes = align256((size_t)csb * (size_t)k * sizeof(float));
    float *psum = reinterpret_cast<float *>(ws);
    float *pmin = reinterpret_cast<float *>(ws + pbytes);
    float *pmax = reinterpret_cast<float *>(ws + 2 * pbytes);
    const bool fp16 = (precision == 2);
    {
        const int64_t cb = (k + 255) / 256;
        hipLaunchKernelGGL(colstats_partial_kernel, dim3((unsigned)csb, (unsigned)(cb < 1024 ? cb : 1024)),
                           dim3(256), 0, stream, x, n, k, ldx, psum, pmin, pmax);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        const int64_t fb = (k + CS_COLS - 1) / CS_COLS;
        hipLaunchKernelGGL(colstats_finish_kernel, dim3((unsigned)(fb < 65536 ? fb : 65536)), dim3(256),
                           0, stream, psum, pmin, pmax, csb, n, k, center);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    {
        int64_t blocks = (n + 3) / 4;
        if (blocks > 65536) blocks = 65536;
        hipLaunchKernelGGL(row_prep_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, x, n, k,
                           ldx, center, norms, fp16 ? scales : nullptr, fp16 ? iscales : nullptr);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    GramArgs args;
    args.x = x;
    args.n = n;
    args.k = k;
    args.ldx = ldx;
    args.terms = terms;
    args.kp = kp;
    args.norms = norms;
    args.iscale = iscales;
    args.out = out;
    args.ldd = ldd;
    args.T = (n + GBM - 1) / GBM;
    args.tau = tau;
    const int64_t S = (args.T + GSUPER - 1) / GSUPER;
    const int64_t ids = S * (S + 1) / 2 * GSUPER * GSUPER;
    const bool vec4 = (k % 4 == 0) && (ldx % 4 == 0) &&
                      ((reinterpret_cast<uintptr_t>(x) & 15u) == 0);
    return fp16 ? launch_split<SPLIT_FP16>(layout, vec4, args, center, scales, ids, stream)
                : launch_split<SPLIT_BF16>(layout, vec4, args, center, scales, ids, stream);
}

}  // namespace gt
