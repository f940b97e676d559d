// qcs_devmath.h -- FP32 exp / log for the fitting kernels.
//
// EXACT: explicit IEEE-754 single-precision operation sequences (Cody-Waite reduced exp with a
// degree-7 Taylor polynomial; log with the range reduction on the bit pattern, f = m - 1 and
// f + f^2 P(f), P of degree 8; the fitting losses' table-driven logs below), so every IEEE
// implementation -- this GPU and the host -- evaluates them bit-identically.  These kernels
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

// ---------------------------------------------------------------------------------------------
// Table-driven logarithm of the multi-emitter fitting losses.  x = m 2^e with m in
// [sqrt(1/2), sqrt(2)) as in qlogf_normal; the top five bits of the reduced mantissa select one
// of 32 intervals j, whose
// table pair is (i_j, t_j) with i_j a float near 1 / (interval centre) and t_j = -log(i_j)
// rounded to float.  i_j was chosen among the floats near the centre's reciprocal so that
// -log(i_j) lies within 3e-4 ulp of t_j (accurate tables); interval 18 holds 1.0 and has the
// exact pair (1, 0).  Then r = m i_j - 1 (one fma, |r| < 2^-6 + 4e-4),
// log(m) = t_j + log1p(r) with log1p(r) = r + r^2 (-1/2 + r (1/3 - r/4)), and
// log(x) = e ln2 + log(m) with ln2 rounded to float in one fma (|log x| >= |e| ln2 / 2, so the
// constant's error stays below 0.07 ulp).  Maximum error 1.25 ulp over all positive normal
// floats.  The table lives in LDS (log_tab_load); the loss folds its data term c into the sum
// before the e ln2 fma (logtab_u).
static __constant__ uint32_t kLogTab[64] = {
    0x3fb31609u, 0xbeabf2bcu, 0x3faf4117u, 0xbea0dfd0u, 0x3fab8713u, 0xbe95dea0u, 0x3fa801dcu, 0xbe8b4090u,
    0x3fa4ad84u, 0xbe8100f9u, 0x3fa168b9u, 0xbe6d7a43u, 0x3f9e4786u, 0xbe596d6du, 0x3f9b5080u, 0xbe46104bu,
    0x3f986dc5u, 0xbe32dbf9u, 0x3f95a207u, 0xbe1fe70du, 0x3f92e464u, 0xbe0cf8f1u, 0x3f905c56u, 0xbdf657d0u,
    0x3f8dd751u, 0xbdd24853u, 0x3f8b658bu, 0xbdaeae36u, 0x3f89230du, 0xbd8d356du, 0x3f86cde9u, 0xbd54260eu,
    0x3f84a776u, 0xbd1249f4u, 0x3f82936au, 0xbca3377au, 0x3f800000u, 0x00000000u, 0x3f7a0ec6u, 0x3cc0651fu,
    0x3f72b5b7u, 0x3d5a5c7cu, 0x3f6bb147u, 0x3da94424u, 0x3f652255u, 0x3de30fcdu, 0x3f5eecc3u, 0x3e0da9b7u,
    0x3f59059fu, 0x3e2924a9u, 0x3f5362ddu, 0x3e441625u, 0x3f4e0ad7u, 0x3e5e4e59u, 0x3f49012au, 0x3e77a781u,
    0x3f442f94u, 0x3e88403au, 0x3f3f8f6fu, 0x3e947796u, 0x3f3b3bffu, 0x3ea02977u, 0x3f3712c1u, 0x3eabab45u,
};
constexpr float kLn2F = 0.693147182464599609375f;   // ln 2 rounded to float

// Copy the table into LDS; every thread of the block calls it (blockDim.x >= 32).
__device__ __forceinline__ void log_tab_load(float2 *lt)
{
    const int l = threadIdx.x;
    if (l < 32) lt[l] = make_float2(__uint_as_float(kLogTab[2 * l]), __uint_as_float(kLogTab[2 * l + 1]));
    __syncthreads();
}

// log(x) + c for positive normal x, where c is small against log x or cancels it
__device__ __forceinline__ float logtab_u(float x, float c, const float2 *lt)
{
    const uint32_t bx = __float_as_uint(x);
    const uint32_t ix = bx - 0x3f3504f3u;
    const float m = __uint_as_float(bx - (ix & 0xff800000u));
    const float2 tv = lt[(ix >> 18) & 31u];
    const float r = __builtin_fmaf(m, tv.x, -1.0f);
    const float r2 = r * r;
    float p = __builtin_fmaf(r, -0.25f, 0.3333333432674407958984375f);
    p = __builtin_fmaf(p, r, -0.5f);
    const float lr = __builtin_fmaf(p, r2, r);
    const float w = (tv.y + lr) + c;
    const float ef = (float)(((int)ix) >> 23);
    return __builtin_fmaf(ef, kLn2F, w);
}

// logtab_u for every x: subnormals are pre-scaled by 2^23; zero gives -inf, negative and NaN
// arguments NaN, +inf gives +inf
__device__ __forceinline__ float logtab_u_any(float x, float c, const float2 *lt)
{
    const bool sub = (x > 0.0f) && (x < 1.17549435082228751e-38f);
    const float xs = sub ? x * 8388608.0f : x;
    const uint32_t bx = __float_as_uint(xs);
    const uint32_t ix = bx - 0x3f3504f3u;
    const float m = __uint_as_float(bx - (ix & 0xff800000u));
    const float2 tv = lt[(ix >> 18) & 31u];
    const float r = __builtin_fmaf(m, tv.x, -1.0f);
    const float r2 = r * r;
    float p = __builtin_fmaf(r, -0.25f, 0.3333333432674407958984375f);
    p = __builtin_fmaf(p, r, -0.5f);
    const float lr = __builtin_fmaf(p, r2, r);
    const float w = (tv.y + lr) + c;
    const float ef = (float)((((int)ix) >> 23) - (sub ? 23 : 0));
    float u = __builtin_fmaf(ef, kLn2F, w);
    u = (x == 0.0f) ? -__builtin_inff() : u;
    u = (x < 0.0f || x != x) ? __builtin_nanf("") : u;
    u = (x == __builtin_inff()) ? x : u;
    return u;
}

// ---------------------------------------------------------------------------------------------
// Binade table log of the one-emitter loss (qcs_logtab.inc, written by logtab_gen.py).  For a
// positive float x with bits b, g = (int)(b - 0x3f368000) >> 16 numbers its interval: binade
// e = g >> 7 of x = m 2^e, m in [0.712890625, 1.42578125), and one of 128 bit-pattern steps of m.
// Binades -10..21, x in [kLogTabBLo, kLogTabBHi), have their own pair (i, t) at g + 1280, with
// t = -log(i) to within 1e-3 ulp (accurate tables) and i near the reciprocal of the interval's
// centre; the interval around 1 has (1, 0).  There log x + c = (t + c) + log1p(r),
// r = fma(x, i, -1), |r| < 2^-8 + 4e-4, log1p(r) = r + r^2 (-1/2 + r / 3): the binade's e ln2
// is part of t, so the loss needs no exponent extraction, conversion or e ln2 FMA, and the byte
// offset of a pair is one add and one shift of the upper half-word (the table is 32 KB).
// Outside that range x is scaled by 2^-e into binade 0 and e ln2 is added in one FMA.
// Max error of log x alone: 1.05 ulp in range.
static __constant__ uint32_t kLogTabB[8192] = {
#include "qcs_logtab.inc"
};
constexpr uint32_t kLogTabBK = 0x3f368000u;
constexpr int kLogTabBN = 4096;               // pairs
constexpr float kLogTabBLo = 0x1.6dp-11f;     // 0.712890625 * 2^-10
constexpr float kLogTabBHi = 0x1.6dp21f;      // 0.712890625 * 2^22

// Copy the 32 KB table into LDS; every thread of the block calls it.
__device__ __forceinline__ void log_tabb_load(float2 *lt)
{
    const uint4 *src = reinterpret_cast<const uint4 *>(kLogTabB);
    uint4 *dst = reinterpret_cast<uint4 *>(lt);
    for (int i = threadIdx.x; i < kLogTabBN / 2; i += blockDim.x) dst[i] = src[i];
    __syncthreads();
}

__device__ __forceinline__ int logtabb_index(uint32_t bx) { return (((int)(bx - kLogTabBK)) >> 16) + 1280; }

__device__ __forceinline__ float logtabb_core(float x, float2 tv, float c)
{
    const float r = __builtin_fmaf(x, tv.x, -1.0f);
    const float r2 = r * r;
    const float p = __builtin_fmaf(r, 0.3333333432674407958984375f, -0.5f);
    const float lr = __builtin_fmaf(p, r2, r);
    return (tv.y + c) + lr;
}

// log(x) + c for x in [kLogTabBLo, kLogTabBHi)
__device__ __forceinline__ float logtabb_u(float x, float c, const float2 *lt)
{
    return logtabb_core(x, lt[logtabb_index(__float_as_uint(x))], c);
}

// logtabb_u for every x: outside the table's range x is scaled into binade 0 (subnormals first
// by 2^23); zero gives -inf, negative and NaN arguments NaN, +inf gives +inf.  The table index
// stays inside binade 0 for every bit pattern outside the range.
__device__ __forceinline__ float logtabb_u_any(float x, float c, const float2 *lt)
{
    const bool inr = (x >= kLogTabBLo) && (x < kLogTabBHi);
    const bool sub = (x > 0.0f) && (x < 1.17549435082228751e-38f);
    const float xs = sub ? x * 8388608.0f : x;
    const uint32_t bx = __float_as_uint(xs);
    const uint32_t ix = bx - kLogTabBK;
    const float xr = inr ? x : __uint_as_float(bx - (ix & 0xff800000u));
    const float w = logtabb_core(xr, lt[logtabb_index(__float_as_uint(xr))], c);
    const float ef = (float)((((int)ix) >> 23) - (sub ? 23 : 0));
    float u = inr ? w : __builtin_fmaf(ef, kLn2F, w);
    u = (x == 0.0f) ? -__builtin_inff() : u;
    u = (x < 0.0f || x != x) ? __builtin_nanf("") : u;
    u = (x == __builtin_inff()) ? x : u;
    return u;
}

__device__ __forceinline__ float fexpf(float x) { return __builtin_amdgcn_exp2f(x * 1.44269504088896341f); }
__device__ __forceinline__ float flogf(float x) { return __builtin_amdgcn_logf(x) * 0.693147180559945309f; }

template <bool FAST> __device__ __forceinline__ float vexp_nonpos(float x) { return FAST ? fexpf(x) : qexpf_nonpos(x); }


}  // namespace qcs
