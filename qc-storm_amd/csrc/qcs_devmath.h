// qcs_devmath.h -- FP32 exp / log for the fitting kernels.
//
// EXACT: explicit IEEE-754 single-precision operation sequences (Cody-Waite reduced exp with a
// degree-7 Taylor polynomial; log with the range reduction on the bit pattern, f = m - 1 and
// f + f^2 P(f), P of degree 8), so every IEEE implementation -- this GPU and the host --
// evaluates them bit-identically.  These kernels
// are compiled with -ffp-contract=off: only the fmaf calls written here are fused.
// FAST: the gfx950 transcendental unit (v_exp_f32 / v_log_f32, base 2).
#pragma once
#include <hip/hip_runtime.h>

namespace qcs {

// exp(x) for x <= 0 or NaN (the loss and weight arguments): no overflow clamp needed
__device__ __forceinline__ float qexpf_nonpos(float x)
{
    const float n = __builtin_rintf(x * 1.44269504088896341f);
    float r = __builtin_fmaf(-n, 0.693145751953125f, x);
    r = __builtin_fmaf(-n, 1.428606765330187e-06f, r);
    float p = 1.98412698412698413e-04f;
    p = __builtin_fmaf(p, r, 1.38888888888888889e-03f);
    p = __builtin_fmaf(p, r, 8.33333333333333333e-03f);
    p = __builtin_fmaf(p, r, 4.16666666666666667e-02f);
    p = __builtin_fmaf(p, r, 1.66666666666666667e-01f);
    p = __builtin_fmaf(p, r, 0.5f);
    p = __builtin_fmaf(p, r, 1.0f);
    p = __builtin_fmaf(p, r, 1.0f);
    const float y = __builtin_ldexpf(p, (int)n);
    return (x < -87.0f) ? 0.0f : y;
}

__device__ __forceinline__ float qexpf(float x)
{
    const float y = qexpf_nonpos(x);
    return (x > 88.0f) ? __builtin_inff() : y;
}

// log(x) for positive normal x (range reduction on the bit pattern)
__device__ __forceinline__ float qlogf_normal(float x)
{
    const uint32_t ix = __float_as_uint(x) - 0x3f3504f3u;
    const int e = ((int)ix) >> 23;
    const float m = __uint_as_float((ix & 0x007fffffu) + 0x3f3504f3u);
    const float f = m - 1.0f;
    float p = -0.07866668701171875f;
    p = __builtin_fmaf(p, f, 0.12821750342845917f);
    p = __builtin_fmaf(p, f, -0.131016343832016f);
    p = __builtin_fmaf(p, f, 0.1418900042772293f);
    p = __builtin_fmaf(p, f, -0.1662759929895401f);
    p = __builtin_fmaf(p, f, 0.2000196874141693f);
    p = __builtin_fmaf(p, f, -0.2500077188014984f);
    p = __builtin_fmaf(p, f, 0.33333322405815125f);
    p = __builtin_fmaf(p, f, -0.4999999701976776f);
    const float f2 = f * f;
    const float r = __builtin_fmaf(p, f2, f);
    const float ef = (float)e;
    return __builtin_fmaf(ef, 0.693145751953125f, __builtin_fmaf(ef, 1.428606765330187e-06f, r));
}

// true when x is a positive normal number (v_cmp_class)
__device__ __forceinline__ bool is_pos_normal(float x) { return __builtin_amdgcn_classf(x, 0x100); }

// log(x) for every x: zero, negative, subnormal, infinite and NaN arguments included
__device__ __forceinline__ float qlogf(float x)
{
    const bool sub = (x > 0.0f) && (x < 1.17549435082228751e-38f);
    const float xs = sub ? x * 8388608.0f : x;
    const uint32_t ix = __float_as_uint(xs) - 0x3f3504f3u;
    const int e = (((int)ix) >> 23) - (sub ? 23 : 0);
    const float m = __uint_as_float((ix & 0x007fffffu) + 0x3f3504f3u);
    const float f = m - 1.0f;
    float p = -0.07866668701171875f;
    p = __builtin_fmaf(p, f, 0.12821750342845917f);
    p = __builtin_fmaf(p, f, -0.131016343832016f);
    p = __builtin_fmaf(p, f, 0.1418900042772293f);
    p = __builtin_fmaf(p, f, -0.1662759929895401f);
    p = __builtin_fmaf(p, f, 0.2000196874141693f);
    p = __builtin_fmaf(p, f, -0.2500077188014984f);
    p = __builtin_fmaf(p, f, 0.33333322405815125f);
    p = __builtin_fmaf(p, f, -0.4999999701976776f);
    const float f2 = f * f;
    const float r = __builtin_fmaf(p, f2, f);
    const float ef = (float)e;
    float y = __builtin_fmaf(ef, 0.693145751953125f, __builtin_fmaf(ef, 1.428606765330187e-06f, r));
    y = (x == 0.0f) ? -__builtin_inff() : y;
    y = (x < 0.0f || x != x) ? __builtin_nanf("") : y;
    y = (x == __builtin_inff()) ? x : y;
    return y;
}

__device__ __forceinline__ float fexpf(float x) { return __builtin_amdgcn_exp2f(x * 1.44269504088896341f); }
__device__ __forceinline__ float flogf(float x) { return __builtin_amdgcn_logf(x) * 0.693147180559945309f; }

template <bool FAST> __device__ __forceinline__ float vexp_nonpos(float x) { return FAST ? fexpf(x) : qexpf_nonpos(x); }
// fast-path log: the caller checks is_pos_normal on the argument and recomputes with qlogf if not
template <bool FAST> __device__ __forceinline__ float vlog_normal(float x) { return FAST ? flogf(x) : qlogf_normal(x); }
template <bool FAST> __device__ __forceinline__ float vlog(float x) { return FAST ? flogf(x) : qlogf(x); }

}  // namespace qcs
