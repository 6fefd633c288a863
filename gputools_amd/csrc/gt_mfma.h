// gt_mfma.h -- split-precision operands of the gfx950 matrix cores, shared
// by the MFMA tile kernels (euclidean Gram form, Pearson correlation).
//
// An FP32 value is split into 16-bit terms whose products, accumulated in
// FP32 by v_mfma_f32_32x32x16_{bf16,f16}, reproduce the FP32 product to
// within the bounds stated at each Split specialisation.
#pragma once

#include <hip/hip_runtime.h>

namespace gt {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x16_t __attribute__((ext_vector_type(16)));

enum { SPLIT_BF16 = 0, SPLIT_FP16 = 1 };

template <int SPLIT>
struct Split;

template <>
struct Split<SPLIT_BF16> {
    typedef bf16x8_t vec;
    static constexpr int NT = 3;      // split terms
    static __device__ __forceinline__ void split(float v, unsigned short *t)
    {
        const __bf16 h = (__bf16)v;
        const float r1 = v - (float)h;     // exact
        const __bf16 m = (__bf16)r1;
        const __bf16 l = (__bf16)(r1 - (float)m);   // r1 - m exact, and exact in BF16
        t[0] = __builtin_bit_cast(unsigned short, h);
        t[1] = __builtin_bit_cast(unsigned short, m);
        t[2] = __builtin_bit_cast(unsigned short, l);
    }
    static __device__ __forceinline__ f32x16_t mfma(vec a, vec b, f32x16_t c)
    {
        return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
    }
};

template <>
struct Split<SPLIT_FP16> {
    typedef f16x8_t vec;
    static constexpr int NT = 2;
    static __device__ __forceinline__ void split(float v, unsigned short *t)
    {
        const _Float16 h = (_Float16)v;
        const _Float16 l = (_Float16)(v - (float)h);
        t[0] = __builtin_bit_cast(unsigned short, h);
        t[1] = __builtin_bit_cast(unsigned short, l);
    }
    static __device__ __forceinline__ f32x16_t mfma(vec a, vec b, f32x16_t c)
    {
        return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
    }
};

// Correctly rounded square root for x >= 2^-100 (normal) or +inf.  The
// hardware v_sqrt_f32 result s is within 1 ulp; the residuals x - s' s of
// its neighbours s' = s -/+ 1 ulp, each exact in one fma, tell whether the
// root lies below s - ulp/2 or above s + ulp/2.  For x = +inf both residuals
// are NaN and s = +inf stays; for x = 0 the result is 0.
__device__ __forceinline__ float sqrt_rn(float x)
{
    const float s = __builtin_amdgcn_sqrtf(x);
    const float sm = __uint_as_float(__float_as_uint(s) - 1u);
    const float sp = __uint_as_float(__float_as_uint(s) + 1u);
    float r = fmaf(-sm, s, x) <= 0.0f ? sm : s;
    r = fmaf(-sp, s, x) > 0.0f ? sp : r;
    return r;
}

}  // namespace gt
