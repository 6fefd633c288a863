// gt_binary.hip -- the 'binary' distance on bit-packed rows.
//
// gputools' binary_kernel_same (inst/cuda/distance.cu:366-415) compares the
// FP32 values element by element and counts in FP32.  Here each row is first
// packed into 32-bit words (bit c of word w = x[32w + c] != 0, one wave ballot
// per 64 features), with the row's number of set bits in one more word, and a
// tile kernel counts popcount(a & b) per pair and word (one AND and one
// v_bcnt_u32_b32, which adds to an accumulator).  From the row counts pa, pb:
//   on in exactly one row = pa + pb - 2 (a & b),  on in either = pa + pb - (a & b).
// The counts are exact integers; the ratio uses the correctly rounded FP32
// divide, which equals the FP32 rounding of the FP64 ratio R computes
// (double rounding of a quotient from binary64 to binary32 is innocuous), so
// the result is bit-exact against stats::dist rounded to FP32.
#include "gt_common.h"
#include "gt_internal.h"

namespace gt {
namespace {

constexpr int BBM = 128;       // points per tile side
constexpr int BBK = 16;        // words per LDS chunk (512 features)
constexpr int BTHREADS = 256;
constexpr int BLD = BBM + 4;
constexpr int BLOADS = BBM * BBK / BTHREADS;

__global__ __launch_bounds__(256) void binary_pack_kernel(const float *__restrict__ x, int64_t n,
                                                          int64_t k, int64_t ldx,
                                                          uint32_t *__restrict__ bits, int64_t ldw,
                                                          int64_t words)
{
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    for (int64_t row = (int64_t)blockIdx.x * 4 + wave; row < n; row += (int64_t)gridDim.x * 4) {
        const float *xr = x + row * ldx;
        uint32_t *br = bits + row * ldw;
        uint32_t total = 0;
        for (int64_t f0 = 0; f0 < k; f0 += 64) {
            const int64_t f = f0 + lane;
            const bool on = (f < k) && (xr[f] != 0.0f);
            const unsigned long long m = __ballot(on);
            total += (uint32_t)__popcll(m);
            const int64_t w = f0 >> 5;
            if (lane == 0) br[w] = (uint32_t)(m & 0xffffffffull);
            if (lane == 1 && w + 1 < words) br[w + 1] = (uint32_t)(m >> 32);
        }
        if (lane == 0) br[words] = total;   // set bits of the row
    }
}

struct BinaryArgs {
    const uint32_t *a;  // column side (i)
    int64_t na;
    const uint32_t *b;  // row side (j)
    int64_t nb;
    int64_t words, ldw; // data words per row; row stride (words + 1: the row's bit count)
    float *out;
    int64_t ldd;
    int64_t tiles_a, tile_offset, tiles_this;
    bool vec;           // ldd and out allow 16-byte stores (pitched layouts)
};

template <int LAYOUT>
__global__ __launch_bounds__(BTHREADS) void binary_tile_kernel(BinaryArgs args)
{
    __shared__ uint32_t lds[2][BBK][BLD];
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
    const int64_t i0 = (int64_t)I * BBM, j0 = (int64_t)J * BBM;

    uint32_t nand[8][8];
#pragma unroll
    for (int r = 0; r < 8; ++r)
#pragma unroll
        for (int c = 0; c < 8; ++c) nand[r][c] = 0;

    for (int64_t wb = 0; wb < args.words; wb += BBK) {
        __syncthreads();
#pragma unroll
        for (int q = 0; q < BLOADS; ++q) {
            const int e = tid + BTHREADS * q;
            const int row = e >> 4, wc = e & 15;
            const int64_t w = wb + wc;
            const int64_t gi = i0 + row, gj = j0 + row;
            lds[0][wc][row] = (gi < args.na && w < args.words) ? args.a[gi * args.ldw + w] : 0u;
            lds[1][wc][row] = (gj < args.nb && w < args.words) ? args.b[gj * args.ldw + w] : 0u;
        }
        __syncthreads();
        const int nw = (args.words - wb) < BBK ? (int)(args.words - wb) : BBK;
#pragma unroll 2
        for (int wc = 0; wc < nw; ++wc) {
            const uint4 c0 = *reinterpret_cast<const uint4 *>(&lds[0][wc][tx * 4]);
            const uint4 c1 = *reinterpret_cast<const uint4 *>(&lds[0][wc][64 + tx * 4]);
            const uint4 r0 = *reinterpret_cast<const uint4 *>(&lds[1][wc][ty * 4]);
            const uint4 r1 = *reinterpret_cast<const uint4 *>(&lds[1][wc][64 + ty * 4]);
            const uint32_t cv[8] = {c0.x, c0.y, c0.z, c0.w, c1.x, c1.y, c1.z, c1.w};
            const uint32_t rv[8] = {r0.x, r0.y, r0.z, r0.w, r1.x, r1.y, r1.z, r1.w};
#pragma unroll
            for (int r = 0; r < 8; ++r)
#pragma unroll
                for (int c = 0; c < 8; ++c) {
                    uint32_t s = nand[r][c];
                    s += (uint32_t)__popc(cv[c] & rv[r]);
                    asm("" : "+v"(s));   // keeps v_bcnt_u32_b32 accumulating
                    nand[r][c] = s;
                }
        }
    }
    uint32_t pc[8], pr[8];   // set bits of the lane's rows on each side
#pragma unroll
    for (int u = 0; u < 8; ++u) {
        const int64_t i = i0 + (u < 4 ? tx * 4 + u : 64 + tx * 4 + (u - 4));
        const int64_t j = j0 + (u < 4 ? ty * 4 + u : 64 + ty * 4 + (u - 4));
        pc[u] = i < args.na ? args.a[i * args.ldw + args.words] : 0u;
        pr[u] = j < args.nb ? args.b[j * args.ldw + args.words] : 0u;
    }

    // Whole tiles of the pitched layouts (cross, and the off-diagonal tiles
    // of the symmetric one) store 16 bytes at a time in both directions.
    if (LAYOUT != OUT_PACKED && args.vec && i0 + BBM <= args.na && j0 + BBM <= args.nb &&
        (LAYOUT == OUT_CROSS || I > J)) {
        float v[8][8];
#pragma unroll
        for (int r = 0; r < 8; ++r)
#pragma unroll
            for (int c = 0; c < 8; ++c) {
                const uint32_t both = pc[c] + pr[r];
                const uint32_t nx = both - 2u * nand[r][c], no = both - nand[r][c];
                v[r][c] = (no == 0) ? 0.0f : (float)nx / (float)no;
            }
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
#pragma unroll
    for (int r = 0; r < 8; ++r) {
        const int64_t j = j0 + (r < 4 ? ty * 4 + r : 64 + ty * 4 + (r - 4));
        if (j >= args.nb) continue;
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            const int64_t i = i0 + (c < 4 ? tx * 4 + c : 64 + tx * 4 + (c - 4));
            if (i >= args.na) continue;
            // R: no bit on in either row -> 0; else (#on in one)/(#on in either)
            const uint32_t both = pc[c] + pr[r];
            const uint32_t nx = both - 2u * nand[r][c], no = both - nand[r][c];
            const float v = (no == 0) ? 0.0f : (float)nx / (float)no;
            if (LAYOUT == OUT_PACKED) {
                if (i > j) args.out[packed_base(args.na, j) + i] = v;
            } else if (LAYOUT == OUT_FULL) {
                if (i > j) {
                    args.out[j * args.ldd + i] = v;
                    args.out[i * args.ldd + j] = v;
                } else if (i == j) {
                    args.out[i * args.ldd + i] = 0.0f;
                }
            } else {
                args.out[j * args.ldd + i] = v;
            }
        }
    }
}

}  // namespace

int64_t binary_words(int64_t k) { return (k + 31) / 32; }

int64_t binary_ldw(int64_t k) { return binary_words(k) + 1; }

hipError_t launch_binary_pack(const float *x, int64_t n, int64_t k, int64_t ldx, uint32_t *bits,
                              int64_t ldw, hipStream_t stream)
{
    if (n <= 0) return hipSuccess;
    int64_t blocks = (n + 3) / 4;
    if (blocks > 16384) blocks = 16384;
    hipLaunchKernelGGL(binary_pack_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, x, n, k,
                       ldx, bits, ldw, binary_words(k));
    return hipGetLastError();
}

hipError_t launch_binary(int layout, const uint32_t *a, int64_t na, const uint32_t *b, int64_t nb,
                         int64_t words, int64_t ldw, float *out, int64_t ldd, hipStream_t stream)
{
    BinaryArgs args;
    args.a = a;
    args.na = na;
    args.b = b;
    args.nb = nb;
    args.words = words;
    args.ldw = ldw;
    args.out = out;
    args.ldd = ldd;
    args.vec = (ldd % 4 == 0) && ((reinterpret_cast<uintptr_t>(out) & 15u) == 0);
    const int64_t ta = (na + BBM - 1) / BBM, tb = (nb + BBM - 1) / BBM;
    args.tiles_a = ta;
    const int64_t total = (layout == OUT_CROSS) ? ta * tb : ta * (ta + 1) / 2;
    const int64_t max_tiles = int64_t(1) << 23;
    for (int64_t off = 0; off < total; off += max_tiles) {
        args.tile_offset = off;
        args.tiles_this = (total - off) < max_tiles ? (total - off) : max_tiles;
        dim3 grid((unsigned)args.tiles_this), block(BTHREADS);
        if (layout == OUT_PACKED)
            hipLaunchKernelGGL(binary_tile_kernel<OUT_PACKED>, grid, block, 0, stream, args);
        else if (layout == OUT_FULL)
            hipLaunchKernelGGL(binary_tile_kernel<OUT_FULL>, grid, block, 0, stream, args);
        else
            hipLaunchKernelGGL(binary_tile_kernel<OUT_CROSS>, grid, block, 0, stream, args);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

}  // namespace gt
