// qcs_devmath.h -- FP32 exp / log for the fitting kernels.
//
// EXACT: explicit IEEE-754 single-precision operation sequences (Cody-Waite reduced exp with a
// degree-7 Taylor polynomial; log via frexp, f = m - 1 and a degree-10 polynomial), so every
// IEEE implementation -- this GPU and the host -- evaluates them bit-identically.  These kernels
// are compiled with -ffp-contract=off: only the fmaf calls written here are fused.
// FAST: the gfx950 transcendental unit (v_exp_f32 / v_log_f32, base 2).
#pragma once
#include <hip/hip_runtime.h>

namespace qcs {

__device__ __forceinline__ float qexpf(float x)
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
    float y = __builtin_ldexpf(p, (int)n);
    y = (x < -87.0f) ? 0.0f : y;
    y = (x > 88.0f) ? __builtin_inff() : y;
    return y;
}

__device__ __forceinline__ float qlogf(float x)
{
    int e;
    float m = __builtin_frexpf(x, &e);
    const bool lo = m < 0.707106781186547524f;
    m = lo ? m + m : m;
    e = lo ? e - 1 : e;
    const float f = m - 1.0f;
    float g = -0.0810227021574974f;
    g = __builtin_fmaf(g, f, 0.1271267831325531f);
    g = __builtin_fmaf(g, f, -0.12972144782543182f);
    g = __builtin_fmaf(g, f, 0.14202114939689636f);
    g = __builtin_fmaf(g, f, -0.166422501206398f);
    g = __builtin_fmaf(g, f, 0.20001471042633057f);
    g = __builtin_fmaf(g, f, -0.2500029504299164f);
    g = __builtin_fmaf(g, f, 0.333333283662796f);
    const float f2 = f * f;
    float r = __builtin_fmaf(g, f2 * f, -0.5f * f2);
    r = r + f;
    const float ef = (float)e;
    float y = __builtin_fmaf(ef, 0.693145751953125f, __builtin_fmaf(ef, 1.428606765330187e-06f, r));
    y = (x == 0.0f) ? -__builtin_inff() : y;
    y = (x < 0.0f) ? __builtin_nanf("") : y;
    y = (x == __builtin_inff()) ? x : y;
    return y;
}

__device__ __forceinline__ float fexpf(float x) { return __builtin_amdgcn_exp2f(x * 1.44269504088896341f); }
__device__ __forceinline__ float flogf(float x) { return __builtin_amdgcn_logf(x) * 0.693147180559945309f; }

template <bool FAST> __device__ __forceinline__ float vexp(float x) { return FAST ? fexpf(x) : qexpf(x); }
template <bool FAST> __device__ __forceinline__ float vlog(float x) { return FAST ? flogf(x) : qlogf(x); }

}  // namespace qcs
