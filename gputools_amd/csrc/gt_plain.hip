// gt_plain.hip -- FP32 all-pairs distance tiles on the gfx950 VALU.
//
// Replaces the six *_kernel_same kernels of inst/cuda/distance.cu (and the
// unreachable cross kernels).  gputools runs one 32-thread block per pair,
// re-reads both rows from global memory for every pair and reduces through
// shared memory (distance.cu:56-69).  Here a 256-thread workgroup owns a
// 128 x 128 tile of pairs: both 128-point row blocks are staged through LDS in
// 16-feature chunks (k-major, so each lane reads float4 runs), and every lane
// keeps an 8 x 8 register micro-tile of pair accumulators.  Each staged
// feature value is reused by 128 pairs.
//
// Summation is two-level: each 16-feature chunk is summed into a partial that
// is then added into the running total, which bounds the rounding error of a
// sum of k positive terms by about (16 + k/16) ulp instead of k ulp.
//
// The same tile also serves the FP32 form of gpuCor's Pearson correlation
// (OP_DOT: the dot product of two standardised variables).
#include "gt_common.h"
#include "gt_internal.h"

#include <math.h>

#include <type_traits>

namespace gt {
namespace {

constexpr int PBM = 128;       // points per tile side
constexpr int PBK = 16;        // features per LDS chunk
constexpr int PTHREADS = 256;  // 4 waves
constexpr int PLD = PBM + 4;   // LDS row length (padding against bank conflicts)
constexpr int PLOADS = PBM * PBK / PTHREADS;  // elements each lane stages per side

struct PlainArgs {
    const float *a;  // column side (index i): self mode = x
    int64_t lda, na;
    const float *b;  // row side (index j): self mode = x
    int64_t ldb, nb;
    int64_t k;
    float *out;
    int64_t ldd;
    int64_t tiles_a;      // cross mode: tiles along A
    int64_t tile_offset;  // first logical tile of this launch
    int64_t tiles_this;   // tiles in this launch
    float p, inv_p, fk;
    int fold;             // chunks summed in `part` before it is added to `acc`
    bool vec;             // ldd and out allow 16-byte stores (pitched layouts)
};

// Canberra chunk modes.  The quotient |a-b| / |a+b| costs an IEEE divide
// (about ten instructions) unless the hardware reciprocal can stand in for it:
// when every staged value of the chunk is 0, non-finite, or of magnitude in
// [2^-100, 2^100], each nonzero finite |a+b| lies in [2^-123, 2^101] (a sum of
// such values is a multiple of 2^-123), so v_rcp_f32 (1 ulp) neither takes
// nor returns a denormal, and |a-b| * rcp(|a+b|) is within 1.5 ulp of the
// quotient with the same Inf (x/0) and NaN (0/0, Inf/Inf) results.
//   CB_CLEAN:  no 0 and no non-finite value -- no term can be NaN, so the
//              per-pair count is the chunk length, added once per chunk;
//   CB_ZEROS:  zeros but no non-finite value -- the only NaN term is 0/0
//              (a == b == 0), which max(q, 0) drops (IEEE maxNum returns the
//              non-NaN operand); the dropped terms are counted per chunk as
//              popcount(za & zb) of the two rows' 16-bit zero masks;
//   CB_RANGED: non-finite values present -- reciprocal + NaN test per term;
//   CB_IEEE:   a tiny or huge finite value present -- IEEE divide.
enum { CB_CLEAN = 0, CB_RANGED = 1, CB_IEEE = 2, CB_ZEROS = 3 };

template <int OP, int MODE = CB_IEEE>
struct Term {
    // Accumulate one feature of one pair into (part, cnt).
    static __device__ __forceinline__ void add(float &part, int &cnt, float a, float b, float p)
    {
        if (OP == OP_EUCLID) {
            const float t = a - b;
            part = fmaf(t, t, part);
        } else if (OP == OP_MAXIMUM) {
            part = fmaxf(part, fabsf(a - b));
        } else if (OP == OP_MANHATTAN) {
            part += fabsf(a - b);
        } else if (OP == OP_CANBERRA && MODE == CB_CLEAN) {
            part = fmaf(fabsf(a - b), __builtin_amdgcn_rcpf(fabsf(a + b)), part);
        } else if (OP == OP_CANBERRA && MODE == CB_ZEROS) {
            part += fmaxf(fabsf(a - b) * __builtin_amdgcn_rcpf(fabsf(a + b)), 0.0f);
        } else if (OP == OP_CANBERRA && MODE == CB_RANGED) {
            const float q = fabsf(a - b) * __builtin_amdgcn_rcpf(fabsf(a + b));
            const bool ok = (q == q);
            part += ok ? q : 0.0f;
            cnt += ok ? 1 : 0;
        } else if (OP == OP_CANBERRA) {
            // R's rule: a term with 0 numerator and 0 denominator (a == b == 0,
            // the only case where the quotient is NaN for finite input) is
            // omitted and counted as missing; a == -b != 0 gives x/0 = Inf.
            const float q = fabsf(a - b) / fabsf(a + b);
            const bool ok = (q == q);
            part += ok ? q : 0.0f;
            cnt += ok ? 1 : 0;
        } else if (OP == OP_MINK3) {
            const float t = fabsf(a - b);
            part = fmaf(t * t, t, part);
        } else if (OP == OP_DOT) {
            part = fmaf(a, b, part);
        } else {  // OP_MINKP: |t|^p = 2^(p log2|t|); log2(0) = -inf gives 0
            const float t = fabsf(a - b);
            part += exp2f(p * log2f(t));
        }
    }

    static __device__ __forceinline__ float finish(float acc, int cnt, float inv_p, float fk)
    {
        if (OP == OP_EUCLID) return sqrtf(acc);
        if (OP == OP_MAXIMUM || OP == OP_MANHATTAN || OP == OP_DOT) return acc;
        if (OP == OP_CANBERRA) {
            if (cnt == 0) return __builtin_nanf("");
            return ((float)cnt == fk) ? acc : acc / ((float)cnt / fk);
        }
        if (OP == OP_MINK3) return cbrtf(acc);
        return powf(acc, inv_p);
    }
};

template <int OP, int LAYOUT>
__global__ __launch_bounds__(PTHREADS) void plain_tile_kernel(PlainArgs args)
{
    __shared__ float lds[2][PBK][PLD];  // [0] column side (i), [1] row side (j)

    const int tid = threadIdx.x;
    const int tx = tid & 15, ty = tid >> 4;
    const int64_t t = args.tile_offset + xcd_remap(blockIdx.x, args.tiles_this);

    int I, J;
    if (LAYOUT == OUT_CROSS) {
        I = (int)(t % args.tiles_a);
        J = (int)(t / args.tiles_a);
    } else {
        tri_tile(t, I, J);
    }
    const int64_t i0 = (int64_t)I * PBM, j0 = (int64_t)J * PBM;
    const int64_t k = args.k;

    // Global -> register staging (prefetched one chunk ahead).
    float sc[PLOADS], sr[PLOADS];
    auto gload = [&](int64_t kb) {
#pragma unroll
        for (int q = 0; q < PLOADS; ++q) {
            const int e = tid + PTHREADS * q;
            const int row = e >> 4, kc = e & 15;
            const int64_t kk = kb + kc;
            const int64_t gi = i0 + row, gj = j0 + row;
            sc[q] = (gi < args.na && kk < k) ? args.a[gi * args.lda + kk] : 0.0f;
            sr[q] = (gj < args.nb && kk < k) ? args.b[gj * args.ldb + kk] : 0.0f;
        }
    };

    float acc[8][8];
    float part[8][8];
    int cnt[8][8];
#pragma unroll
    for (int r = 0; r < 8; ++r)
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            acc[r][c] = 0.0f;
            part[r][c] = 0.0f;
            cnt[r][c] = 0;
        }

    // One chunk of PBK features from LDS into the micro-tile.
    auto chunk = [&](auto mode) {
        constexpr int MODE = decltype(mode)::value;
#pragma unroll 4
        for (int kc = 0; kc < PBK; ++kc) {
            const float4 c0 = *reinterpret_cast<const float4 *>(&lds[0][kc][tx * 4]);
            const float4 c1 = *reinterpret_cast<const float4 *>(&lds[0][kc][64 + tx * 4]);
            const float4 r0 = *reinterpret_cast<const float4 *>(&lds[1][kc][ty * 4]);
            const float4 r1 = *reinterpret_cast<const float4 *>(&lds[1][kc][64 + ty * 4]);
            const float cv[8] = {c0.x, c0.y, c0.z, c0.w, c1.x, c1.y, c1.z, c1.w};
            const float rv[8] = {r0.x, r0.y, r0.z, r0.w, r1.x, r1.y, r1.z, r1.w};
#pragma unroll
            for (int r = 0; r < 8; ++r)
#pragma unroll
                for (int c = 0; c < 8; ++c)
                    Term<OP, MODE>::add(part[r][c], cnt[r][c], cv[c], rv[r], args.p);
        }
    };

    __shared__ int cb_flags[PTHREADS / 64];  // canberra: per-wave chunk classification
    __shared__ uint32_t cb_zero[2][PBM];     // canberra CB_ZEROS: per-row zero masks
    int clean_terms = 0;                     // canberra: features of CB_CLEAN / CB_ZEROS chunks

    gload(0);
    int folded = 0;   // chunks in `part` since the last fold
    for (int64_t kb = 0; kb < k; kb += PBK) {
        __syncthreads();
        bool unclean = false, unranged = false, nonfin = false;
#pragma unroll
        for (int q = 0; q < PLOADS; ++q) {
            const int e = tid + PTHREADS * q;
            const int row = e >> 4, kc = e & 15;
            lds[0][kc][row] = sc[q];
            lds[1][kc][row] = sr[q];
            if (OP == OP_CANBERRA) {
#pragma unroll
                for (int s = 0; s < 2; ++s) {
                    const float m = fabsf(s ? sr[q] : sc[q]);
                    const bool mid = (m >= 0x1p-100f) && (m <= 0x1p100f);
                    unclean |= !mid;
                    unranged |= !mid && m != 0.0f && m < __builtin_inff();
                    nonfin |= !(m < __builtin_inff());
                }
            }
        }
        if (OP == OP_CANBERRA) {
            const bool wu = __builtin_amdgcn_ballot_w64(unclean) != 0;
            const bool wr = __builtin_amdgcn_ballot_w64(unranged) != 0;
            const bool wn = __builtin_amdgcn_ballot_w64(nonfin) != 0;
            if ((tid & 63) == 0) cb_flags[tid >> 6] = (wu ? 1 : 0) | (wr ? 2 : 0) | (wn ? 4 : 0);
        }
        __syncthreads();
        if (kb + PBK < k) gload(kb + PBK);

        if (OP == OP_CANBERRA) {
            const int4 f = *reinterpret_cast<const int4 *>(cb_flags);
            const int bits = __builtin_amdgcn_readfirstlane(f.x | f.y | f.z | f.w);
            if (bits == 0) {
                chunk(std::integral_constant<int, CB_CLEAN>{});
                clean_terms += PBK;
            } else if (bits == 1) {
                // zero masks of the 2 x 128 staged rows, one row per thread
                {
                    const int side = tid >> 7, row = tid & (PBM - 1);
                    uint32_t zm = 0;
#pragma unroll
                    for (int kc = 0; kc < PBK; ++kc) zm |= (lds[side][kc][row] == 0.0f ? 1u : 0u) << kc;
                    cb_zero[side][row] = zm;
                }
                __syncthreads();
                chunk(std::integral_constant<int, CB_ZEROS>{});
                clean_terms += PBK;
                uint32_t zc[8], zr[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    zc[u] = cb_zero[0][u < 4 ? tx * 4 + u : 64 + tx * 4 + (u - 4)];
                    zr[u] = cb_zero[1][u < 4 ? ty * 4 + u : 64 + ty * 4 + (u - 4)];
                }
#pragma unroll
                for (int r = 0; r < 8; ++r)
#pragma unroll
                    for (int c = 0; c < 8; ++c) cnt[r][c] -= __builtin_popcount(zc[c] & zr[r]);
            } else if ((bits & 2) == 0) {
                chunk(std::integral_constant<int, CB_RANGED>{});
            } else {
                chunk(std::integral_constant<int, CB_IEEE>{});
            }
        } else {
            chunk(std::integral_constant<int, CB_IEEE>{});
        }
        // Two-level sum: `part` collects `fold` chunks (about sqrt(k) features),
        // then joins the total, so both levels add about sqrt(k / 16) terms.
        if (OP != OP_MAXIMUM && (++folded == args.fold || kb + PBK >= k)) {
            folded = 0;
#pragma unroll
            for (int r = 0; r < 8; ++r)
#pragma unroll
                for (int c = 0; c < 8; ++c) {
                    acc[r][c] += part[r][c];
                    part[r][c] = 0.0f;
                }
        }
    }

    // Whole tiles of the pitched layouts (cross, and the off-diagonal tiles
    // of the symmetric one) store 16 bytes at a time: the lane's i runs
    // tx*4.. and 64 + tx*4.. along a row j, and its j runs ty*4.. and
    // 64 + ty*4.. along a mirrored row i.
    if (LAYOUT != OUT_PACKED && args.vec && i0 + PBM <= args.na && j0 + PBM <= args.nb &&
        (LAYOUT == OUT_CROSS || I > J)) {
        float v[8][8];
#pragma unroll
        for (int r = 0; r < 8; ++r)
#pragma unroll
            for (int c = 0; c < 8; ++c)
                v[r][c] = Term<OP>::finish((OP == OP_MAXIMUM) ? part[r][c] : acc[r][c],
                                           cnt[r][c] + clean_terms, args.inv_p, args.fk);
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            const int64_t j = j0 + (r < 4 ? ty * 4 + r : 64 + ty * 4 + (r - 4));
#pragma unroll
            for (int h = 0; h < 2; ++h)
                *reinterpret_cast<float4 *>(args.out + j * args.ldd + i0 + h * 64 + tx * 4) =
                    make_float4(v[r][4 * h], v[r][4 * h + 1], v[r][4 * h + 2], v[r][4 * h + 3]);
        }
        if (LAYOUT == OUT_FULL) {
#pragma unroll
            for (int c = 0; c < 8; ++c) {
                const int64_t i = i0 + (c < 4 ? tx * 4 + c : 64 + tx * 4 + (c - 4));
#pragma unroll
                for (int h = 0; h < 2; ++h)
                    *reinterpret_cast<float4 *>(args.out + i * args.ldd + j0 + h * 64 + ty * 4) =
                        make_float4(v[4 * h][c], v[4 * h + 1][c], v[4 * h + 2][c], v[4 * h + 3][c]);
            }
        }
        return;
    }

    if (LAYOUT == OUT_PACKED) {
        // Packed rows start at arbitrary offsets, so the values go out as
        // whole runs of a row instead: 32 rows of the tile at a time are
        // transposed through LDS (the staging array), then each wave writes
        // rows with its 64 lanes on 64 consecutive i (two 256-byte runs per
        // row).  The lane holding rows ty*4 + r (r < 4) and 64 + ty*4 + r - 4
        // takes part in groups 0-1 and 2-3 respectively.
        float *tr = &lds[0][0][0];   // 32 x TLD floats
        constexpr int TLD = PBM + 1;
        static_assert(32 * TLD * 4 <= (int)sizeof(lds), "transpose buffer");
        const int wv = tid >> 6, ln = tid & 63;
#pragma unroll
        for (int grp = 0; grp < 4; ++grp) {
            __syncthreads();
            const int rbase = (grp & 1) * 32;
            if (ty * 4 >= rbase && ty * 4 < rbase + 32) {
#pragma unroll
                for (int r4 = 0; r4 < 4; ++r4) {
                    const int r = (grp < 2 ? 0 : 4) + r4;
                    float *trow = tr + (ty * 4 + r4 - rbase) * TLD;
#pragma unroll
                    for (int c = 0; c < 8; ++c) {
                        const float total = (OP == OP_MAXIMUM) ? part[r][c] : acc[r][c];
                        trow[c < 4 ? tx * 4 + c : 64 + tx * 4 + (c - 4)] =
                            Term<OP>::finish(total, cnt[r][c] + clean_terms, args.inv_p, args.fk);
                    }
                }
            }
            __syncthreads();
            const int64_t jg = j0 + grp * 32;
#pragma unroll
            for (int rr = 0; rr < PTHREADS / 64 * 2; ++rr) {
                const int lr = wv + (PTHREADS / 64) * rr;
                const int64_t j = jg + lr;
                if (j >= args.nb) continue;
                float *orow = args.out + packed_base(args.na, j);
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    const int64_t i = i0 + h * 64 + ln;
                    if (i > j && i < args.na) orow[i] = tr[lr * TLD + h * 64 + ln];
                }
            }
        }
        return;
    }
#pragma unroll
    for (int r = 0; r < 8; ++r) {
        const int64_t j = j0 + (r < 4 ? ty * 4 + r : 64 + ty * 4 + (r - 4));
        if (j >= args.nb) continue;
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            const int64_t i = i0 + (c < 4 ? tx * 4 + c : 64 + tx * 4 + (c - 4));
            if (i >= args.na) continue;
            const float total = (OP == OP_MAXIMUM) ? part[r][c] : acc[r][c];
            if (LAYOUT == OUT_PACKED) {
                if (i > j)
                    args.out[packed_base(args.na, j) + i] =
                        Term<OP>::finish(total, cnt[r][c] + clean_terms, args.inv_p, args.fk);
            } else if (LAYOUT == OUT_FULL) {
                if (i > j) {
                    const float v = Term<OP>::finish(total, cnt[r][c] + clean_terms, args.inv_p, args.fk);
                    args.out[j * args.ldd + i] = v;
                    args.out[i * args.ldd + j] = v;
                } else if (i == j) {
                    // a distance is 0 there; the Pearson dot product is z . z
                    args.out[i * args.ldd + i] =
                        (OP == OP_DOT) ? Term<OP>::finish(total, cnt[r][c] + clean_terms, args.inv_p, args.fk)
                                       : 0.0f;
                }
            } else {
                args.out[j * args.ldd + i] = Term<OP>::finish(total, cnt[r][c] + clean_terms, args.inv_p, args.fk);
            }
        }
    }
}

template <int OP>
hipError_t launch_op(int layout, PlainArgs &args, int64_t total, hipStream_t stream)
{
    // Keep each launch's thread count below 2^32.
    const int64_t max_tiles = int64_t(1) << 23;
    for (int64_t off = 0; off < total; off += max_tiles) {
        args.tile_offset = off;
        args.tiles_this = (total - off) < max_tiles ? (total - off) : max_tiles;
        dim3 grid((unsigned)args.tiles_this), block(PTHREADS);
        if (layout == OUT_PACKED)
            hipLaunchKernelGGL((plain_tile_kernel<OP, OUT_PACKED>), grid, block, 0, stream, args);
        else if (layout == OUT_FULL)
            hipLaunchKernelGGL((plain_tile_kernel<OP, OUT_FULL>), grid, block, 0, stream, args);
        else
            hipLaunchKernelGGL((plain_tile_kernel<OP, OUT_CROSS>), grid, block, 0, stream, args);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

}  // namespace

hipError_t launch_plain(int op, int layout, const float *a, int64_t na, int64_t lda,
                        const float *b, int64_t nb, int64_t ldb, int64_t k, float p,
                        float *out, int64_t ldd, hipStream_t stream)
{
    PlainArgs args;
    args.a = a;
    args.lda = lda;
    args.na = na;
    args.b = b;
    args.ldb = ldb;
    args.nb = nb;
    args.k = k;
    args.out = out;
    args.ldd = ldd;
    args.vec = (ldd % 4 == 0) && ((reinterpret_cast<uintptr_t>(out) & 15u) == 0);
    args.p = p;
    args.inv_p = 1.0f / p;
    args.fk = (float)k;
    {
        // chunks per fold: about sqrt(chunks), so that the partial and the
        // total each add about sqrt(k / 16) values
        const int64_t chunks = (k + PBK - 1) / PBK;
        int f = 1;
        while ((int64_t)(f + 1) * (f + 1) <= chunks) ++f;
        args.fold = f;
    }
    const int64_t ta = (na + PBM - 1) / PBM, tb = (nb + PBM - 1) / PBM;
    int64_t total;
    if (layout == OUT_CROSS) {
        args.tiles_a = ta;
        total = ta * tb;
    } else {
        args.tiles_a = ta;
        total = ta * (ta + 1) / 2;
    }
    if (total == 0) return hipSuccess;
    switch (op) {
    case OP_EUCLID: return launch_op<OP_EUCLID>(layout, args, total, stream);
    case OP_MAXIMUM: return launch_op<OP_MAXIMUM>(layout, args, total, stream);
    case OP_MANHATTAN: return launch_op<OP_MANHATTAN>(layout, args, total, stream);
    case OP_CANBERRA: return launch_op<OP_CANBERRA>(layout, args, total, stream);
    case OP_MINK3: return launch_op<OP_MINK3>(layout, args, total, stream);
    case OP_MINKP: return launch_op<OP_MINKP>(layout, args, total, stream);
    case OP_DOT: return launch_op<OP_DOT>(layout, args, total, stream);
    default: return hipErrorInvalidValue;
    }
}

// ---------------------------------------------------------------------------
// Fills: k == 0 gives NA for every pair (R: no usable column -> NA_REAL).
// ---------------------------------------------------------------------------
namespace {
__global__ void fill_kernel(float *out, int64_t count, float value)
{
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < count;
         e += (int64_t)gridDim.x * blockDim.x)
        out[e] = value;
}

__global__ void fill_full_kernel(float *out, int64_t n, int64_t ldd, float offdiag)
{
    const int64_t total = n * n;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = e / n, c = e % n;
        out[r * ldd + c] = (r == c) ? 0.0f : offdiag;
    }
}
}  // namespace

hipError_t launch_fill(float *out, int64_t count, float value, hipStream_t stream)
{
    if (count <= 0) return hipSuccess;
    int64_t blocks = (count + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(fill_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, out, count, value);
    return hipGetLastError();
}

hipError_t launch_fill_full(float *out, int64_t n, int64_t ldd, float offdiag, hipStream_t stream)
{
    if (n <= 0) return hipSuccess;
    int64_t blocks = (n * n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(fill_full_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, out, n, ldd,
                       offdiag);
    return hipGetLastError();
}

}  // namespace gt
