// gt_binary.hip -- the 'binary' distance on bit-packed rows.
//
// gputools' binary_kernel_same (inst/cuda/distance.cu:366-415) compares the
// FP32 values element by element and counts in FP32.  Here each row is first
// packed into 32-bit words (bit c of word w = x[32w + c] != 0, one wave ballot
// per 64 features), with the row's number of set bits in one more word, and a
// tile kernel counts popcount(a & b) per pair and word (one AND and one
// v_bcnt_u32_b32, which adds to an accumulator).  From the row counts pa, pb:
//   on in exactly one row = pa + pb - 2 (a & b),  on in either = pa + pb - (a & b).
// Packed requests on one point set of up to SMALL_DIV_MAX features take the
// matrix cores instead: one 0/1 byte per feature and v_mfma_i32_32x32x32_i8
// give the same exact count of features on in both rows (binary_mfma_kernel).
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

// The same value with y = RN(1/no) read from a table (lut[0] = 0: no = 0
// implies nx = 0, and the sequence then gives +0).  RN(1/no) is one of the
// reciprocals tests/exhaustive/small_div.c checks.
__device__ __forceinline__ float binary_value_lut(uint32_t pa_plus_pb, uint32_t both, const float *lut)
{
    const uint32_t no = pa_plus_pb - both, nx = no - both;
    const float y = lut[no];
    const float a = (float)nx, b = (float)no;
    const float q0 = __fmul_rn(a, y);
    const float r = fmaf(-q0, b, a);
    return fmaf(r, y, q0);
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

// ---------------------------------------------------------------------------
// Packed whole-matrix requests: the counts on the matrix cores.  Each row is
// one 0/1 byte per feature (zero-padded to a multiple of 32) and its count of
// ones; v_mfma_i32_32x32x32_i8 sums the products of two rows' bytes exactly,
// which is the number of features on in both.  The tiles are 256 x 256 (two
// workgroups of 8 waves) or 128 x 128 (4 waves); chunks of 32 features (32 B
// per row) reach LDS through global_load_lds into a ring of NSTAGE stages,
// 16-byte halves swapped by ((R >> 2) ^ (R >> 3)) & 1 so that the operand
// reads of consecutive rows hit distinct banks.
// ---------------------------------------------------------------------------
typedef int bin_i32x4_t __attribute__((ext_vector_type(4)));
typedef int bin_i32x16_t __attribute__((ext_vector_type(16)));

constexpr int QBK = 32;                  // features per chunk (bytes per row)
constexpr int QSUPER = 4;                // super-tile side in tiles
// Tile side: 256 x 256 pairs for large n; 128 x 128 (four workgroups per
// CU) while the 256 schedule has fewer than QSMALL_ROUNDS rounds of 256 CUs
// (measured with the first matrix-core form, 64 x 64 pairs per wave:
// 65536 x 256 2.73 vs 3.07 ms, 8192 x 256 0.049 vs 0.043 ms).
constexpr int64_t QSMALL_ROUNDS = 4;

// Tile geometry.  Square tiles of QB x QB pairs (B rows i, A rows j) are
// split along j into QB / QA workgroups of QA columns each.  A workgroup
// computes its QA columns against QB + QX rows: the QX extra rows let it
// finish the 128-byte line that its last row of every column ends in (the
// line-ownership rule below).  Wave (wj, wi) holds A rows 32 wj .. 32 wj + 31
// against B rows 128 wi .. 128 wi + 159: five 32 x 32 blocks, the fifth
// shared with the next wave (or the extra rows).  Two 8-wave workgroups per
// 256 tile ran as fast as one of 16 waves (2.52 vs 2.51 ms at 65536 x 256)
// with less LDS per workgroup.
constexpr int QX = 32;
constexpr int QXB = 5;                          // B blocks per wave
// Reciprocal table entries (no = 0 .. 2 kb) when 2 kb < QLUT: an LDS read
// instead of v_rcp_f32 (a quarter-rate instruction) and the no = 0 select.
constexpr int QLUT = 2048;
template <int QA, int QB>
struct QTile {
    static constexpr int WJ = QA / 32;             // waves along j
    static constexpr int WI = QB / 128;            // waves along i
    static constexpr int NW = WJ * WI;
    static constexpr int THREADS = NW * 64;
    static constexpr int PARTS = QB / QA;          // workgroups per square tile
    static constexpr int GA = QA / 32;             // 32-row load groups: A side,
    static constexpr int G = GA + (QB + QX) / 32;  // then B side
    static constexpr int GMAX = (G + NW - 1) / NW; // groups of wave 0
    static constexpr int TERM = QA * QBK;          // the A side of one chunk
    static constexpr int STAGE = G * 32 * QBK;
    static constexpr int NSTAGE = 3;               // ring stages
};

// s_waitcnt vmcnt(v) for a wave-uniform v (larger v: waits for everything)
__device__ __forceinline__ void wait_vmcnt(int v)
{
    switch (v) {
    case 1: asm volatile("s_waitcnt vmcnt(1)" ::: "memory"); break;
    case 2: asm volatile("s_waitcnt vmcnt(2)" ::: "memory"); break;
    case 3: asm volatile("s_waitcnt vmcnt(3)" ::: "memory"); break;
    case 4: asm volatile("s_waitcnt vmcnt(4)" ::: "memory"); break;
    case 5: asm volatile("s_waitcnt vmcnt(5)" ::: "memory"); break;
    case 6: asm volatile("s_waitcnt vmcnt(6)" ::: "memory"); break;
    case 7: asm volatile("s_waitcnt vmcnt(7)" ::: "memory"); break;
    case 8: asm volatile("s_waitcnt vmcnt(8)" ::: "memory"); break;
    case 9: asm volatile("s_waitcnt vmcnt(9)" ::: "memory"); break;
    case 10: asm volatile("s_waitcnt vmcnt(10)" ::: "memory"); break;
    case 11: asm volatile("s_waitcnt vmcnt(11)" ::: "memory"); break;
    case 12: asm volatile("s_waitcnt vmcnt(12)" ::: "memory"); break;
    default: asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); break;
    }
}

int64_t bytes_ld(int64_t k) { return (k + QBK - 1) / QBK * QBK; }

// One wave per row: bytes (x != 0) of 4 features per lane, and the row's
// number of ones.
__global__ __launch_bounds__(256) void binary_pack8_kernel(const float *__restrict__ x, int64_t n,
                                                           int64_t k, int64_t ldx, int64_t kb,
                                                           uint32_t *__restrict__ ind,
                                                           uint32_t *__restrict__ cnt)
{
    const int lane = threadIdx.x & 63;
    for (int64_t row = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6); row < n;
         row += (int64_t)gridDim.x * 4) {
        const float *xr = x + row * ldx;
        uint32_t *ir = ind + row * (kb / 4);
        uint32_t ones = 0;
        for (int64_t f0 = 0; f0 < kb; f0 += 256) {
            const int64_t f = f0 + 4 * lane;
            if (f >= kb) break;
            uint32_t w = 0;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const bool on = (f + e < k) && (xr[f + e] != 0.0f);
                w |= (on ? 1u : 0u) << (8 * e);
                ones += on ? 1u : 0u;
            }
            ir[f / 4] = w;
        }
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) ones += __shfl_xor(ones, d, 64);
        if (lane == 0) cnt[row] = ones;
    }
}

struct BinMfmaArgs {
    const unsigned char *ind;   // n x kb bytes
    const uint32_t *cnt;        // ones per row
    int64_t n, kb;
    float *out;                 // packed
    int64_t T;                  // tiles per side
    int64_t id_offset, ids_this;
    int64_t stagger_below;      // odd workgroups below this id start late (first round)
};

// Line ownership.  Packed column j (the pairs (i, j), i > j) is one run of
// the output.  A tile that wrote exactly its rows i0 .. i0 + QB - 1 of it
// would share the 128-byte lines at both ends with the tiles above and
// below, which write them at other times (often through another XCD's L2),
// so those lines would reach HBM as partial writes.  Here the 32-float line
// holding row i of column j belongs to the tile whose rows hold
// s(i) = max(j + 1, the first row of that line): tile I writes rows
// [i0 + e_j, i0 + QB + e_j) of column j, where e_j is the distance from row
// i0 to the next line start (line_gap), except that the tile holding j + 1
// starts at j + 1 and a tile below j + 1 writes nothing of the column.  The
// QX extra B rows cover the overhang.  Only a line that holds the end of one
// column and the start of the next is still written by two tiles.
//
// Inside a tile, wave wi writes rows [128 wi + e_j, 128 wi + 128 + e_j) of
// its columns as four whole lines: each half-wave rotates its five blocks'
// values down by e_j lanes (ds_bpermute), so one store instruction writes
// one aligned line per half-wave.
__device__ __forceinline__ unsigned line_gap(uint64_t float_index)
{
    return (unsigned)(0u - (unsigned)float_index) & 31u;
}

__device__ __forceinline__ float lane_fetch(float v, int src_byte)
{
    return __int_as_float(__builtin_amdgcn_ds_bpermute(src_byte, __float_as_int(v)));
}

// Only called with at most SMALL_DIV_MAX features: the quotient is
// small_quotient's.
template <int QA, int QB, bool LUT>
__global__ __launch_bounds__((QTile<QA, QB>::THREADS), (1024 / QTile<QA, QB>::THREADS)) void binary_mfma_kernel(
    BinMfmaArgs args)
{
    typedef QTile<QA, QB> Q;
    constexpr int QTERM = Q::TERM, QSTAGE = Q::STAGE, NW = Q::NW, NS = Q::NSTAGE, G = Q::G, GA = Q::GA;
    static_assert((NS - 2) * Q::GMAX <= 12, "wait_vmcnt covers the loads in flight");
    __shared__ __attribute__((aligned(1024))) unsigned char smem[NS * QSTAGE + (QA + QB + QX) * 4];
    uint32_t *cA = reinterpret_cast<uint32_t *>(smem + NS * QSTAGE);
    uint32_t *cB = cA + QA;
    __shared__ float lut[LUT ? QLUT : 1];
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wj = wave % Q::WJ, wi = wave / Q::WJ;   // A rows (j) 32 wj.., B rows (i) 128 wi..

    const int64_t gid = args.id_offset + xcd_remap(blockIdx.x, args.ids_this);
    // Two 8-wave workgroups share a CU and have the same work, so if they
    // start together they stay in step, and both store at once and then both
    // run the matrix cores.  The odd workgroups of the first round start
    // about 0.4 of a tile late, which keeps the pairs offset afterwards
    // (65536 x 256, A/B on one box: 2.12 vs 2.30 ms).
    if ((blockIdx.x & 1) && (int64_t)blockIdx.x < args.stagger_below) {
        __builtin_amdgcn_s_sleep(127);
        __builtin_amdgcn_s_sleep(127);
    }
    const int part = (int)(gid % Q::PARTS);
    const int64_t id = gid / Q::PARTS;
    const int64_t st = id / (QSUPER * QSUPER);
    const int w = (int)(id % (QSUPER * QSUPER));
    int SI, SJ;
    tri_tile(st, SI, SJ);
    const int64_t I = (int64_t)SI * QSUPER + w / QSUPER;
    const int64_t J = (int64_t)SJ * QSUPER + w % QSUPER;
    if (I >= args.T || J > I) return;
    const int64_t i0 = I * QB, j0 = J * QB + part * QA, n = args.n;
    if (j0 >= n) return;
    for (int t = tid; t < QA + QB + QX; t += Q::THREADS) {
        if (t < QA)
            cA[t] = j0 + t < n ? args.cnt[j0 + t] : 0u;
        else
            cB[t - QA] = i0 + t - QA < n ? args.cnt[i0 + t - QA] : 0u;
    }
    if (LUT)
        for (int t = tid; t <= 2 * args.kb; t += Q::THREADS) lut[t] = t ? 1.0f / (float)t : 0.0f;
    auto value = [&](uint32_t pa_plus_pb, uint32_t both) {
        return LUT ? binary_value_lut(pa_plus_pb, both, lut) : binary_value<true>(pa_plus_pb, both);
    };

    // 32-row load groups: g < GA are A rows 32 g.., the rest B rows
    // 32 (g - GA)..; group g lands at byte 1024 g of a stage.  Wave w fills
    // groups w, w + NW, ... below G.  Rows past n are clamped to n - 1 (they
    // only feed pairs never stored).
    typedef __attribute__((address_space(3))) void lds_void_t;
    typedef const __attribute__((address_space(1))) void gbl_void_t;
    const unsigned char *gsrc[Q::GMAX];
    unsigned char *ldst[Q::GMAX];
#pragma unroll
    for (int q = 0; q < Q::GMAX; ++q) {
        const int g = wave + q * NW < G ? wave + q * NW : G - 1;
        const int rr = lane >> 1;
        const int half = (lane & 1) ^ (((rr >> 2) ^ (rr >> 3)) & 1);
        const int64_t r = (g < GA ? j0 + g * 32 : i0 + (g - GA) * 32) + rr;
        gsrc[q] = args.ind + (r < n ? r : n - 1) * args.kb + half * 16;
        ldst[q] = smem + g * 1024;
    }
    auto issue = [&](int chunk) {
#pragma unroll
        for (int q = 0; q < Q::GMAX; ++q)
            if (wave + q * NW < G)
                __builtin_amdgcn_global_load_lds((gbl_void_t *)(gsrc[q] + (int64_t)chunk * QBK),
                                                 (lds_void_t *)(ldst[q] + (chunk % NS) * QSTAGE), 16, 0,
                                                 0);
    };
    const int lrow = lane & 31;
    const int roff = lrow * QBK + (((lane >> 5) ^ (((lrow >> 2) ^ (lrow >> 3)) & 1)) * 16);

    bin_i32x16_t acc[QXB];
#pragma unroll
    for (int b = 0; b < QXB; ++b)
#pragma unroll
        for (int r = 0; r < 16; ++r) acc[b][r] = 0;

    const int nchunks = (int)(args.kb / QBK);
    const int lpc = (G - wave + NW - 1) / NW;   // this wave's loads per chunk
    for (int c = 0; c < NS - 1 && c < nchunks; ++c) issue(c);
    for (int c = 0; c < nchunks; ++c) {
        // chunks c .. c + ahead are in flight; chunk c landed when at most
        // the later ones' loads of this wave are outstanding (vector memory
        // loads complete in order)
        const int ahead = (nchunks - 1 - c) < NS - 2 ? (nchunks - 1 - c) : NS - 2;
        wait_vmcnt(ahead * lpc);
        __builtin_amdgcn_s_barrier();
        asm volatile("" ::: "memory");
        // the stage read in the previous iteration, which every wave finished
        if (c + NS - 1 < nchunks) issue(c + NS - 1);
        const unsigned char *sg = smem + (c % NS) * QSTAGE + roff;
        const bin_i32x4_t av = *reinterpret_cast<const bin_i32x4_t *>(sg + wj * 32 * QBK);
#pragma unroll
        for (int b = 0; b < QXB; ++b) {
            const bin_i32x4_t bv =
                *reinterpret_cast<const bin_i32x4_t *>(sg + QTERM + (wi * 128 + b * 32) * QBK);
            acc[b] = __builtin_amdgcn_mfma_i32_32x32x32_i8(av, bv, acc[b], 0, 0, 0);
        }
    }
    __syncthreads();   // cA, cB and lut visible

    // 32x32 accumulator layout: element r of lane l is A row 8(r/4) +
    // 4(l/32) + r%4 (j) and B row l%32 (i): lanes 0..31 hold 32 consecutive
    // i of one j, block b rows 128 wi + 32 b.
    const int lx = lane & 31;
    const bool hi = (lane >> 5) != 0;
    const int lib = wi * 128 + lx;
    uint32_t pb[QXB];
#pragma unroll
    for (int b = 0; b < QXB; ++b) pb[b] = cB[lib + b * 32];
    const uint64_t fout = reinterpret_cast<uintptr_t>(args.out) >> 2;
    if (I > J && i0 + QB + QX <= n) {
        // every computed pair valid, i > j: row offsets from the wave's
        // first row jw fit in 32 bits (n < 2^24)
        const int64_t jw = j0 + wj * 32;
        float *const outw = args.out + packed_base(n, jw) + i0;
        const unsigned fw = (unsigned)fout + (unsigned)(packed_base(n, jw) + i0);
        const unsigned cw = (unsigned)(n - 1 - jw);
        // column j = i0 - 1 (the last column of the tile row above) starts
        // at i0: wave wi = 0 also writes its rows before the first line start
        const int lastc = (j0 + QA == i0 && wi == 0) ? QA - 1 - wj * 32 - (hi ? 4 : 0) : -1;
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const unsigned dc = 8 * (r >> 2) + (r & 3);
            const unsigned base_u = dc * cw - dc * (dc + 1) / 2;
            const unsigned step_u = 4 * cw - 4 * dc - 10;
            const unsigned rc = base_u + (hi ? step_u : 0u);   // column j at row i0
            const unsigned e = (0u - (fw + rc)) & 31u;
            const uint32_t pa = cA[wj * 32 + (int)dc + (hi ? 4 : 0)];
            float v[QXB];
#pragma unroll
            for (int b = 0; b < QXB; ++b) v[b] = value(pa + pb[b], (uint32_t)acc[b][r]);
            // byte offsets from the wave-uniform outw stay in 32 bits
            char *const ob = reinterpret_cast<char *>(outw);
            if ((int)dc == lastc && (unsigned)lx < e)
                *reinterpret_cast<float *>(ob + 4u * (rc + (unsigned)lx)) = v[0];
            // line m, lane lx: row 128 wi + e + 32 m + lx, held by lane
            // (lx + e) % 32 in block m, or in block m + 1 when lx + e >= 32;
            // so source lane l sends block m + 1 if l < e, else block m
            const int src = ((lane & 32) | (int)(((unsigned)lx + e) & 31u)) << 2;
            const bool sends_next = (unsigned)lx < e;
            const unsigned lb = rc + (unsigned)(wi * 128) + e + (unsigned)lx;
#pragma unroll
            for (int m = 0; m < 4; ++m)
                *reinterpret_cast<float *>(ob + 4u * (lb + 32u * m)) =
                    lane_fetch(sends_next ? v[m + 1] : v[m], src);
        }
        return;
    }
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int lj = wj * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
        const int64_t j = j0 + lj;
        if (j >= n || j + 1 >= i0 + QB) continue;
        const int64_t pbj = packed_base(n, j);
        const int64_t e = line_gap(fout + (uint64_t)(pbj + i0));
        const int64_t lo = j + 1 >= i0 ? j + 1 : i0 + e;
        const int64_t hi_ = i0 + QB + e < n ? i0 + QB + e : n;
        const uint32_t pa = cA[lj];
#pragma unroll
        for (int b = 0; b < QXB; ++b) {
            // the last block is the next wave's first, except for the last wave
            if (b == QXB - 1 && wi != Q::WI - 1) continue;
            const int64_t i = i0 + lib + b * 32;
            if (i >= lo && i < hi_)
                args.out[pbj + i] = value(pa + pb[b], (uint32_t)acc[b][r]);
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

template <int QA, int QB, bool LUT>
hipError_t launch_mfma_tiles(BinMfmaArgs &args, hipStream_t stream)
{
    typedef QTile<QA, QB> Q;
    args.T = (args.n + QB - 1) / QB;
    const int64_t S = (args.T + QSUPER - 1) / QSUPER;
    const int64_t ids = S * (S + 1) / 2 * QSUPER * QSUPER * Q::PARTS;
    args.stagger_below = 0;
    if (Q::PARTS == 2) {   // two workgroups per CU: stagger the first round
        int dev = 0, cus = 0;
        hipError_t e = hipGetDevice(&dev);
        if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        if (e != hipSuccess) return e;
        args.stagger_below = 2 * (int64_t)cus;
    }
    const int64_t max_ids = int64_t(1) << 23;   // a multiple of PARTS
    for (int64_t off = 0; off < ids; off += max_ids) {
        args.id_offset = off;
        args.ids_this = (ids - off) < max_ids ? (ids - off) : max_ids;
        hipLaunchKernelGGL((binary_mfma_kernel<QA, QB, LUT>), dim3((unsigned)args.ids_this), dim3(Q::THREADS),
                           0, stream, args);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

}  // namespace

int64_t binary_words(int64_t k) { return (k + 31) / 32; }

// Workspace of a binary call on one point set: the bit rows (with their
// counts), then, up to SMALL_DIV_MAX features, the byte rows and counts of
// the matrix-core form (a quarter of the input's size).
size_t binary_workspace_bytes(int64_t n, int64_t k)
{
    auto a256 = [](size_t v) { return (v + 255) / 256 * 256; };
    size_t b = a256((size_t)n * (size_t)binary_ldw(k) * sizeof(uint32_t));
    if (k <= SMALL_DIV_MAX) b += a256((size_t)n * (size_t)bytes_ld(k)) + a256((size_t)n * sizeof(uint32_t));
    return b;
}

hipError_t launch_binary_self(int layout, const float *x, int64_t n, int64_t k, int64_t ldx,
                              void *workspace, float *out, int64_t ldd, hipStream_t stream)
{
    auto a256 = [](size_t v) { return (v + 255) / 256 * 256; };
    char *ws = static_cast<char *>(workspace);
    uint32_t *bits = reinterpret_cast<uint32_t *>(ws);
    const int64_t words = binary_words(k), ldw = binary_ldw(k);
    if (layout != OUT_PACKED || n >= (int64_t(1) << 24) || k > SMALL_DIV_MAX) {
        hipError_t e = launch_binary_pack(x, n, k, ldx, bits, ldw, stream);
        if (e != hipSuccess) return e;
        return launch_binary(layout, bits, n, bits, n, words, ldw, out, ldd, stream);
    }
    const int64_t kb = bytes_ld(k);
    ws += a256((size_t)n * (size_t)ldw * sizeof(uint32_t));
    unsigned char *ind = reinterpret_cast<unsigned char *>(ws);
    ws += a256((size_t)n * (size_t)kb);
    uint32_t *cnt = reinterpret_cast<uint32_t *>(ws);
    {
        int64_t blocks = (n + 3) / 4;
        if (blocks > 65536) blocks = 65536;
        hipLaunchKernelGGL(binary_pack8_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, x, n, k,
                           ldx, kb, reinterpret_cast<uint32_t *>(ind), cnt);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    BinMfmaArgs args;
    args.ind = ind;
    args.cnt = cnt;
    args.n = n;
    args.kb = kb;
    args.out = out;
    const bool lut = 2 * kb < QLUT;
    const int64_t t256 = (n + 255) / 256;
    if (t256 * (t256 + 1) / 2 < QSMALL_ROUNDS * 256)
        return lut ? launch_mfma_tiles<128, 128, true>(args, stream)
                   : launch_mfma_tiles<128, 128, false>(args, stream);
    return lut ? launch_mfma_tiles<128, 256, true>(args, stream) : launch_mfma_tiles<128, 256, false>(args, stream);
}

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
