// gt_batched.hip -- all-pairs distances of many independent point sets in one
// persistent launch.
//
// gputools computes one matrix per call (distance() at src/distance.cpp:259-316
// with an n x n grid of one-pair blocks); a caller with thousands of small,
// unevenly sized sets pays one launch (and, in the reference, one module load)
// per set.  Here a host-built schedule lists every lower-triangular tile
// (set, I, J) of every set, sorted by decreasing cost; a persistent grid of
// workgroups strides through it (workgroup b takes entries b, b + G, ...),
// so the launch needs no counters or resets and can be captured in a HIP graph
// and replayed.  Each tile is a BM x BM block of pairs computed with exact FP32
// differences through LDS (the plain-path arithmetic, two-level summation over
// 16-feature chunks, the next chunk's loads in registers while one is
// computed), written in R 'dist' order into the set's own packed output.
//
// Tile shapes (gt_batched_tile_count): 1 = 32 x 32, 2 = 64 x 64, 3 = 128 x 128
// pairs per tile, all with 256 threads; the autotuner picks one per workload.
#include "gt_common.h"
#include "gt_internal.h"

namespace gt {
namespace {

constexpr int BTHREADS = 256;
constexpr int BKC = 16;         // features per LDS chunk

struct BatchedArgs {
    const float *x;
    const int64_t *x_off, *sizes, *dims, *out_off;
    float *out;
    const int64_t *schedule;   // (set << 32) | (I << 16) | J
    int64_t count, nsets;
};

template <int OP>
__device__ __forceinline__ void bterm(float &part, float a, float b)
{
    if (OP == OP_EUCLID) {
        const float t = a - b;
        part = fmaf(t, t, part);
    } else if (OP == OP_MANHATTAN) {
        part += fabsf(a - b);
    } else {
        part = fmaxf(part, fabsf(a - b));
    }
}

// Index within a BM-point tile side of element m of thread t's TM points:
// for TM = 8 two runs of four (t * 4 and BM / 2 + t * 4), so that the 16
// threads of a side read consecutive 16 B pieces of an LDS row; for smaller
// TM one run of TM.
template <int BM, int TM>
__device__ __forceinline__ int side_index(int t, int m)
{
    if (TM == 8) return (m >> 2) * (BM / 2) + t * 4 + (m & 3);
    return t * TM + m;
}

template <int OP, int BM>
__global__ __launch_bounds__(BTHREADS) void batched_kernel(BatchedArgs args)
{
    constexpr int TPS = 16;             // threads per tile side
    constexpr int TM = BM / TPS;        // pairs per thread per side
    constexpr int LD = BM + 4;
    constexpr int LOADS = BM * BKC / BTHREADS;
    __shared__ __attribute__((aligned(16))) float as_[BKC][LD];
    __shared__ __attribute__((aligned(16))) float bs_[BKC][LD];

    const int tid = threadIdx.x;
    const int ty = tid / TPS, tx = tid % TPS;

    for (int64_t wi = blockIdx.x; wi < args.count; wi += gridDim.x) {
        const int64_t e = args.schedule[wi];
        const int64_t s = e >> 32;
        const int I = (int)((e >> 16) & 0xffff), J = (int)(e & 0xffff);
        if (s < 0 || s >= args.nsets) continue;
        const int64_t n = args.sizes[s], k = args.dims[s];
        if ((int64_t)I * BM >= n || J > I) continue;   // not a tile of this set
        const float *x = args.x + args.x_off[s];
        float *out = args.out + args.out_off[s];
        const int64_t i0 = (int64_t)I * BM, j0 = (int64_t)J * BM;

        // global -> register staging, one chunk ahead of the LDS chunk
        float sa[LOADS], sb[LOADS];
        auto gload = [&](int64_t kb) {
#pragma unroll
            for (int q = 0; q < LOADS; ++q) {
                const int u = tid + BTHREADS * q;
                const int row = u / BKC, kc = u % BKC;
                const int64_t kk = kb + kc;
                const int64_t ri = i0 + row, rj = j0 + row;
                sa[q] = (rj < n && kk < k) ? x[rj * k + kk] : 0.0f;   // A side: j
                sb[q] = (ri < n && kk < k) ? x[ri * k + kk] : 0.0f;   // B side: i
            }
        };

        float acc[TM][TM], part[TM][TM];
#pragma unroll
        for (int a = 0; a < TM; ++a)
#pragma unroll
            for (int b = 0; b < TM; ++b) acc[a][b] = part[a][b] = 0.0f;

        gload(0);
        for (int64_t kb = 0; kb < k; kb += BKC) {
            __syncthreads();
#pragma unroll
            for (int q = 0; q < LOADS; ++q) {
                const int u = tid + BTHREADS * q;
                const int row = u / BKC, kc = u % BKC;
                as_[kc][row] = sa[q];
                bs_[kc][row] = sb[q];
            }
            __syncthreads();
            if (kb + BKC < k) gload(kb + BKC);
            // the chunk's real features only (a set with d = 8 fills half)
            const int nk = (k - kb) < BKC ? (int)(k - kb) : BKC;
#pragma unroll 4
            for (int kc = 0; kc < nk; ++kc) {
                float av[TM], bv[TM];
#pragma unroll
                for (int m = 0; m < TM; ++m) {
                    av[m] = as_[kc][side_index<BM, TM>(ty, m)];
                    bv[m] = bs_[kc][side_index<BM, TM>(tx, m)];
                }
#pragma unroll
                for (int a = 0; a < TM; ++a)
#pragma unroll
                    for (int b = 0; b < TM; ++b) bterm<OP>(part[a][b], av[a], bv[b]);
            }
            if (OP != OP_MAXIMUM) {
#pragma unroll
                for (int a = 0; a < TM; ++a)
#pragma unroll
                    for (int b = 0; b < TM; ++b) {
                        acc[a][b] += part[a][b];
                        part[a][b] = 0.0f;
                    }
            }
        }
#pragma unroll
        for (int a = 0; a < TM; ++a) {
            const int64_t j = j0 + side_index<BM, TM>(ty, a);
            if (j >= n) continue;
            const int64_t base = packed_base(n, j);
#pragma unroll
            for (int b = 0; b < TM; ++b) {
                const int64_t i = i0 + side_index<BM, TM>(tx, b);
                if (i < n && i > j) {
                    const float v = (OP == OP_MAXIMUM) ? part[a][b]
                                    : (OP == OP_EUCLID) ? sqrtf(acc[a][b]) : acc[a][b];
                    out[base + i] = v;
                }
            }
        }
    }
}

template <int BM>
hipError_t launch_bm(int op, const BatchedArgs &args, int grid, hipStream_t stream)
{
    if (op == OP_EUCLID)
        hipLaunchKernelGGL((batched_kernel<OP_EUCLID, BM>), dim3(grid), dim3(BTHREADS), 0, stream, args);
    else if (op == OP_MANHATTAN)
        hipLaunchKernelGGL((batched_kernel<OP_MANHATTAN, BM>), dim3(grid), dim3(BTHREADS), 0, stream,
                           args);
    else
        hipLaunchKernelGGL((batched_kernel<OP_MAXIMUM, BM>), dim3(grid), dim3(BTHREADS), 0, stream, args);
    return hipGetLastError();
}

template <int BM>
int resident_blocks(int op)
{
    int per_cu = 0;
    hipError_t e;
    if (op == OP_EUCLID)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, batched_kernel<OP_EUCLID, BM>, BTHREADS, 0);
    else if (op == OP_MANHATTAN)
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, batched_kernel<OP_MANHATTAN, BM>, BTHREADS, 0);
    else
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, batched_kernel<OP_MAXIMUM, BM>, BTHREADS, 0);
    if (e != hipSuccess || per_cu < 1) per_cu = 1;
    return per_cu;
}

}  // namespace

int batched_tile_side(int tile)
{
    switch (tile) {
    case 1: return 32;
    case 2: return 64;
    case 3: return 128;
    default: return 0;
    }
}

hipError_t launch_batched(int op, int tile, const float *x, const int64_t *x_off,
                          const int64_t *sizes, const int64_t *dims, const int64_t *out_off,
                          float *out, const int64_t *schedule, int64_t count, int64_t nsets,
                          int grid, hipStream_t stream)
{
    if (count <= 0) return hipSuccess;
    BatchedArgs args;
    args.x = x;
    args.x_off = x_off;
    args.sizes = sizes;
    args.dims = dims;
    args.out_off = out_off;
    args.out = out;
    args.schedule = schedule;
    args.count = count;
    args.nsets = nsets;
    if (tile == 1) return launch_bm<32>(op, args, grid, stream);
    if (tile == 2) return launch_bm<64>(op, args, grid, stream);
    return launch_bm<128>(op, args, grid, stream);
}

int batched_resident_blocks(int op, int tile)
{
    if (tile == 1) return resident_blocks<32>(op);
    if (tile == 2) return resident_blocks<64>(op);
    return resident_blocks<128>(op);
}

}  // namespace gt
