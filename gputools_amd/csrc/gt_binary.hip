// gt_binary.hip -- the 'binary' distance on bit-packed rows.
//
// gputools' binary_kernel_same (inst/cuda/distance.cu:366-415) compares the
// FP32 values element by element and counts in FP32.  Here each row is first
// packed into 32-bit words (bit c of word w = x[32w + c] != 0, one wave ballot
// per 64 features), with the row's number of set bits in one more word, and a
// tile kernel counts popcount(a & b) per pair and word (one AND and one
// v_bcnt_u32_b32, which adds to an accumulator).  From the row counts pa, pb:
//   on in exactly one row = pa + pb - 2 (a & b),  on in either = pa + pb - (a & b).
// The counts are exact integers; their ratio is rounded correctly to FP32,
// which equals the FP32 rounding of the FP64 ratio R computes (double
// rounding of a quotient from binary64 to binary32 is innocuous), so the
// result is bit-exact against stats::dist rounded to FP32.  Up to
// SMALL_DIV_MAX features the quotient takes a reciprocal and two fused
// corrections instead of the IEEE divide sequence (small_quotient).
#include "gt_common.h"
#include "gt_internal.h"

namespace gt {
namespace {

constexpr int BBM = 128;       // points per tile side
constexpr int BBK = 16;        // words per LDS chunk (512 features)
constexpr int BTHREADS = 256;
constexpr int BLD = BBM + 4;
constexpr int BLOADS = BBM * BBK / BTHREADS;

// Largest count for which small_quotient is exact (checked exhaustively).
constexpr int64_t SMALL_DIV_MAX = 131072;

// a / b rounded correctly, for integers 0 <= a <= b <= SMALL_DIV_MAX held
// exactly as floats: y = v_rcp_f32(b) (within 1 ulp of 1/b), q0 = a y, the
// remainder a - q0 b is exact in one fma, and q0 + r y rounds to the
// correctly rounded quotient.  tests/exhaustive/small_div.c checks every
// such pair against each float within one ulp of 1/b.
__device__ __forceinline__ float small_quotient(float a, float b)
{
    const float y = __builtin_amdgcn_rcpf(b);
    const float q0 = __fmul_rn(a, y);
    const float r = fmaf(-q0, b, a);
    return fmaf(r, y, q0);
}

// R's binary distance from the counts: no bit on in either row -> 0, else
// (#on in exactly one) / (#on in either).
template <bool FASTDIV>
__device__ __forceinline__ float binary_value(uint32_t pa_plus_pb, uint32_t both)
{
    const uint32_t nx = pa_plus_pb - 2u * both, no = pa_plus_pb - both;
    const float q = FASTDIV ? small_quotient((float)nx, (float)no) : (float)nx / (float)no;
    return no == 0 ? 0.0f : q;
}

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

// One 128 x 128 tile of pairs per workgroup.  Wave w holds tile rows (B
// side, j) w * 32 .. w * 32 + 31 and lane l columns (A side, i) l and l + 64:
// 32 x 2 counts per lane.  A row's words are the same for the whole wave
// (broadcast LDS reads), and a store instruction writes 64 consecutive
// values of one output row, from a wave-uniform row base.
template <int LAYOUT, bool FASTDIV>
__global__ __launch_bounds__(BTHREADS) void binary_tile_kernel(BinaryArgs args)
{
    __shared__ __attribute__((aligned(16))) uint32_t lds[2][BBK][BLD];
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int64_t t = args.tile_offset + xcd_remap(blockIdx.x, args.tiles_this);
    int I, J;
    if (LAYOUT == OUT_CROSS) {
        I = (int)(t % args.tiles_a);
        J = (int)(t / args.tiles_a);
    } else {
        tri_tile(t, I, J);
    }
    const int64_t i0 = (int64_t)I * BBM, j0 = (int64_t)J * BBM;
    const int rw = wave * 32;   // first tile row of this wave

    uint32_t nand[32][2];
#pragma unroll
    for (int r = 0; r < 32; ++r) nand[r][0] = nand[r][1] = 0;

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
        for (int wc = 0; wc < nw; ++wc) {
            const uint32_t c0 = lds[0][wc][lane], c1 = lds[0][wc][64 + lane];
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const uint4 rq = *reinterpret_cast<const uint4 *>(&lds[1][wc][rw + 4 * q]);
                const uint32_t rv[4] = {rq.x, rq.y, rq.z, rq.w};
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    uint32_t s0 = nand[4 * q + u][0], s1 = nand[4 * q + u][1];
                    s0 += (uint32_t)__popc(c0 & rv[u]);
                    s1 += (uint32_t)__popc(c1 & rv[u]);
                    asm("" : "+v"(s0), "+v"(s1));   // keeps v_bcnt_u32_b32 accumulating
                    nand[4 * q + u][0] = s0;
                    nand[4 * q + u][1] = s1;
                }
            }
        }
    }
    // set bits of the lane's two columns; the rows' counts are wave-uniform
    uint32_t pc[2];
#pragma unroll
    for (int c = 0; c < 2; ++c) {
        const int64_t i = i0 + lane + 64 * c;
        pc[c] = i < args.na ? args.a[i * args.ldw + args.words] : 0u;
    }
    const uint32_t *const prow = args.b + args.words;
    const bool whole = i0 + BBM <= args.na && j0 + BBM <= args.nb;

    if (LAYOUT == OUT_PACKED && I > J && whole) {
        // every pair valid and i > j
#pragma unroll
        for (int r = 0; r < 32; ++r) {
            const int64_t j = j0 + rw + r;
            const uint32_t pr = prow[j * args.ldw];
            float *rowp = args.out + packed_base(args.na, j) + i0;
#pragma unroll
            for (int c = 0; c < 2; ++c) rowp[lane + 64 * c] = binary_value<FASTDIV>(pc[c] + pr, nand[r][c]);
        }
        return;
    }
    if (LAYOUT != OUT_PACKED && whole && (LAYOUT == OUT_CROSS || I > J)) {
        // direct rows: 64 consecutive values per store instruction
        float v[32][2];
#pragma unroll
        for (int r = 0; r < 32; ++r) {
            const int64_t j = j0 + rw + r;
            const uint32_t pr = prow[j * args.ldw];
#pragma unroll
            for (int c = 0; c < 2; ++c) {
                v[r][c] = binary_value<FASTDIV>(pc[c] + pr, nand[r][c]);
                args.out[j * args.ldd + i0 + lane + 64 * c] = v[r][c];
            }
        }
        if (LAYOUT == OUT_FULL) {
            // mirrored: a lane's 32 rows are 128 consecutive bytes of row i
#pragma unroll
            for (int c = 0; c < 2; ++c) {
                float *dst = args.out + (i0 + lane + 64 * c) * args.ldd + j0 + rw;
                if (args.vec) {
#pragma unroll
                    for (int q = 0; q < 8; ++q)
                        *reinterpret_cast<float4 *>(dst + 4 * q) =
                            make_float4(v[4 * q][c], v[4 * q + 1][c], v[4 * q + 2][c], v[4 * q + 3][c]);
                } else {
#pragma unroll
                    for (int r = 0; r < 32; ++r) dst[r] = v[r][c];
                }
            }
        }
        return;
    }
#pragma unroll
    for (int r = 0; r < 32; ++r) {
        const int64_t j = j0 + rw + r;
        if (j >= args.nb) break;
        const uint32_t pr = prow[j * args.ldw];
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            const int64_t i = i0 + lane + 64 * c;
            if (i >= args.na) continue;
            const float v = binary_value<FASTDIV>(pc[c] + pr, nand[r][c]);
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

template <bool FASTDIV>
void launch_layout(int layout, dim3 grid, dim3 block, hipStream_t stream, const BinaryArgs &args)
{
    if (layout == OUT_PACKED)
        hipLaunchKernelGGL((binary_tile_kernel<OUT_PACKED, FASTDIV>), grid, block, 0, stream, args);
    else if (layout == OUT_FULL)
        hipLaunchKernelGGL((binary_tile_kernel<OUT_FULL, FASTDIV>), grid, block, 0, stream, args);
    else
        hipLaunchKernelGGL((binary_tile_kernel<OUT_CROSS, FASTDIV>), grid, block, 0, stream, args);
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
        // the counts are at most 32 x words
        if (words * 32 <= SMALL_DIV_MAX)
            launch_layout<true>(layout, grid, block, stream, args);
        else
            launch_layout<false>(layout, grid, block, stream, args);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

}  // namespace gt
