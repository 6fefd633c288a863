// gt_calib.hip -- peak-rate probes behind the measured figures quoted next to
// the published MI355X peaks: register-only loops of independent FP32 FMAs,
// of 32-bit AND + popcount-accumulate, and of BF16 MFMAs, a streaming copy
// and a write-only fill.  Built into its own library (libgt_calib.so); the distance library
// does not use it.  Each probe returns the kernel time of one launch in ms
// (hipEvent pair on the given stream), and the caller derives the rate from
// the work it states.
#include <hip/hip_runtime.h>
#include <stdint.h>

typedef __bf16 calib_bf16x8_t __attribute__((ext_vector_type(8)));
typedef float calib_f32x16_t __attribute__((ext_vector_type(16)));

namespace {

constexpr int CHAINS = 16;   // independent dependency chains per lane

// 16 independent FMA chains per lane: 2 flop x 16 per iteration.
__global__ __launch_bounds__(256) void fma_loop(float a, float b, int iters, float *out)
{
    float v[CHAINS];
#pragma unroll
    for (int i = 0; i < CHAINS; ++i) v[i] = (float)(threadIdx.x + i) * 1e-3f;
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int i = 0; i < CHAINS; ++i) v[i] = fmaf(v[i], a, b);
    }
    float s = 0.0f;
#pragma unroll
    for (int i = 0; i < CHAINS; ++i) s += v[i];
    out[(int64_t)blockIdx.x * blockDim.x + threadIdx.x] = s;
}

// 16 independent AND + popcount-accumulate chains per lane: 2 integer lane
// operations x 16 per iteration.  The operand is made opaque each iteration
// so the popcounts cannot be hoisted out of the loop.
__global__ __launch_bounds__(256) void popc_loop(uint32_t y0, int iters, uint32_t *out)
{
    uint32_t x[CHAINS], acc[CHAINS];
#pragma unroll
    for (int i = 0; i < CHAINS; ++i) {
        x[i] = (threadIdx.x + 1u) * 2654435761u + (uint32_t)i * 40503u;
        acc[i] = 0u;
    }
    uint32_t y = y0;
    for (int it = 0; it < iters; ++it) {
        asm volatile("" : "+v"(y));
#pragma unroll
        for (int i = 0; i < CHAINS; ++i) {
            uint32_t s = acc[i] + (uint32_t)__popc(x[i] & y);
            asm("" : "+v"(s));   // keeps v_bcnt_u32_b32 accumulating
            acc[i] = s;
        }
    }
    uint32_t s = 0u;
#pragma unroll
    for (int i = 0; i < CHAINS; ++i) s ^= acc[i];
    out[(int64_t)blockIdx.x * blockDim.x + threadIdx.x] = s;
}

// 8 independent chains of v_mfma_f32_32x32x16_bf16 per wave: 32768 flop
// each.
__global__ __launch_bounds__(256) void mfma_loop(int iters, float *out)
{
    calib_bf16x8_t a, b;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        a[i] = (__bf16)((float)((threadIdx.x + i) & 7) * 0.125f);
        b[i] = (__bf16)((float)((threadIdx.x * 3 + i) & 7) * 0.0625f);
    }
    calib_f32x16_t acc[8];
#pragma unroll
    for (int j = 0; j < 8; ++j)
#pragma unroll
        for (int r = 0; r < 16; ++r) acc[j][r] = 0.0f;
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int j = 0; j < 8; ++j) acc[j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc[j], 0, 0, 0);
    }
    float s = 0.0f;
#pragma unroll
    for (int j = 0; j < 8; ++j)
#pragma unroll
        for (int r = 0; r < 16; ++r) s += acc[j][r];
    out[(int64_t)blockIdx.x * blockDim.x + threadIdx.x] = s;
}

__global__ __launch_bounds__(256) void copy_kernel(const float4 *__restrict__ src,
                                                   float4 *__restrict__ dst, int64_t n4)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4;
         i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

__global__ __launch_bounds__(256) void fill_kernel(float4 *__restrict__ dst, int64_t n4, float v)
{
    const float4 w = make_float4(v, v, v, v);
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4;
         i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = w;
}

template <typename Launch>
int timed(hipStream_t stream, float *ms, Launch launch)
{
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess) return -1;
    if (hipEventCreate(&e1) != hipSuccess) {
        hipEventDestroy(e0);
        return -1;
    }
    hipEventRecord(e0, stream);
    launch();
    hipError_t err = hipGetLastError();
    hipEventRecord(e1, stream);
    if (err == hipSuccess) err = hipEventSynchronize(e1);
    if (err == hipSuccess) err = hipEventElapsedTime(ms, e0, e1);
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    return err == hipSuccess ? 0 : -1;
}

}  // namespace

extern "C" {

// blocks x 256 threads; out holds blocks * 256 floats.  Work: blocks * 256 *
// iters * 32 flop.
int gt_calib_fma(int blocks, int iters, float *out, hipStream_t stream, float *ms)
{
    return timed(stream, ms, [&] {
        hipLaunchKernelGGL(fma_loop, dim3(blocks), dim3(256), 0, stream, 0.999999f, 1e-6f, iters, out);
    });
}

// Work: blocks * 256 * iters * 32 integer lane operations.
int gt_calib_popc(int blocks, int iters, uint32_t *out, hipStream_t stream, float *ms)
{
    return timed(stream, ms, [&] {
        hipLaunchKernelGGL(popc_loop, dim3(blocks), dim3(256), 0, stream, 0x9e3779b9u, iters, out);
    });
}

// Work: blocks * 4 waves * iters * 8 * 32768 flop.
int gt_calib_mfma(int blocks, int iters, float *out, hipStream_t stream, float *ms)
{
    return timed(stream, ms, [&] {
        hipLaunchKernelGGL(mfma_loop, dim3(blocks), dim3(256), 0, stream, iters, out);
    });
}

// Copies n4 float4 values; traffic 32 * n4 bytes.
int gt_calib_copy(const void *src, void *dst, int64_t n4, hipStream_t stream, float *ms)
{
    return timed(stream, ms, [&] {
        hipLaunchKernelGGL(copy_kernel, dim3(8192), dim3(256), 0, stream,
                           static_cast<const float4 *>(src), static_cast<float4 *>(dst), n4);
    });
}

// Writes n4 float4 values; traffic 16 * n4 bytes (write only).
int gt_calib_fill(void *dst, int64_t n4, hipStream_t stream, float *ms)
{
    return timed(stream, ms, [&] {
        hipLaunchKernelGGL(fill_kernel, dim3(8192), dim3(256), 0, stream, static_cast<float4 *>(dst), n4,
                           1.0f);
    });
}

}  // extern "C"
