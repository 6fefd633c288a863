// gt_common.h -- device helpers shared by the gfx950 distance kernels.
//
// Tile scheduling: all-pairs work is cut into square tiles of BM points per
// side.  The self-distance case only needs the tiles on or below the diagonal
// (tile row I >= tile column J), numbered in row order: t = I(I+1)/2 + J.
// gputools launches an n x n grid with one 32-thread block per pair and lets
// the upper half exit (src/distance.cpp:215-216, inst/cuda/distance.cu:51);
// here the grid is 1-D over the T(T+1)/2 lower tiles only.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gt {

// Number of XCDs on MI355X.  Workgroups are dispatched round-robin over the
// XCDs (workgroup b runs on XCD b % 8); each XCD has its own L2.
constexpr int kNumXcd = 8;

// Remap a hardware workgroup id to a logical tile id so that every XCD owns a
// contiguous range of tile ids (neighbouring tiles share input rows, which
// then stay in that XCD's L2).  A bijection on [0, total).
__device__ __forceinline__ int64_t xcd_remap(int64_t b, int64_t total)
{
    const int64_t q = total / kNumXcd, r = total % kNumXcd;
    const int64_t x = b % kNumXcd, local = b / kNumXcd;
    const int64_t start = x * q + (x < r ? x : r);
    return start + local;
}

// Lower-triangular tile id -> (I, J), J <= I.
__device__ __forceinline__ void tri_tile(int64_t t, int &I, int &J)
{
    // I = floor((sqrt(8t+1)-1)/2), corrected for rounding.
    double s = sqrt(8.0 * (double)t + 1.0);
    int64_t i = (int64_t)((s - 1.0) * 0.5);
    while ((i + 1) * (i + 2) / 2 <= t) ++i;
    while (i * (i + 1) / 2 > t) --i;
    I = (int)i;
    J = (int)(t - i * (i + 1) / 2);
}

// Start of column j in R 'dist' order, minus (j + 1): the pair (i, j), i > j,
// is at packed_base(n, j) + i.
__device__ __forceinline__ int64_t packed_base(int64_t n, int64_t j)
{
    return j * n - j * (j + 1) / 2 - j - 1;
}

}  // namespace gt
