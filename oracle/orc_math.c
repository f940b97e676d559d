/*
 * orc_math.c -- deterministic FP32 exp and natural log used by the oracle (TEST INFRASTRUCTURE).
 *
 * The reference evaluates its loss with the CUDA fast intrinsics __expf/__logf
 * (DLL/bfgsMLE_Loc_LD/bfgsMLE_GS2D_core.h:501-503).  Those are hardware approximations that
 * cannot be reproduced off NVIDIA hardware, so the oracle states the exponential and logarithm
 * as explicit FP32 operation sequences written from textbook methods:
 *
 *   exp(x): Cody-Waite range reduction x = n ln2 + r with ln2 split in a high and a low part,
 *           |r| <= ln2/2, then the degree-7 Taylor polynomial of e^r (truncation error
 *           r^8/8! < 5.1e-9 relative), scaled by 2^n with ldexp.
 *   log(x): x = m 2^e with m in [sqrt(1/2), sqrt(2)) taken from the IEEE bit pattern
 *           (subnormals pre-scaled by 2^23), f = m - 1 exact (Sterbenz),
 *           log(1+f) = f + f^2 P(f) with P a degree-8 polynomial fitted by least squares on
 *           Chebyshev nodes of [sqrt(1/2)-1, sqrt(2)-1] (max error 0.82 ulp of log1p),
 *           plus e ln2 (split constant).
 *   table log (the fit loss's logarithm, orc_tlogf / orc_tlog_u): the same reduction to
 *           m in [sqrt(1/2), sqrt(2)); the top five bits of the reduced mantissa select one of
 *           32 intervals, each with a pair (i_j, t_j): i_j a float near 1/(interval centre),
 *           t_j = -log(i_j) rounded to float, i_j picked among nearby floats so that -log(i_j)
 *           is within 3e-4 ulp of t_j (accurate tables; the interval holding 1 has (1, 0)).
 *           r = fma(m, i_j, -1), log1p(r) = r + r^2 (-1/2 + r (1/3 - r/4)) (|r| < 2^-6 + 4e-4,
 *           truncation < 2e-10), log x = fma(e, ln2_f, t_j + log1p(r)).  The loss adds its
 *           data term c before the last fma: orc_tlog_u(x, c) = log x + c.  Max error of
 *           orc_tlogf 1.25 ulp over all positive normal floats.  The multi-emitter losses use it.
 *   binade table log (the one-emitter losses, orc_tlogb_u): g = (int)(b - 0x3f368000) >> 16 of
 *           the bit pattern b numbers the binade e and one of 128 intervals of x = m 2^e,
 *           m in [0.712890625, 1.42578125); for e in -10..21 the pair (i, t) at g + 1280 carries
 *           e ln2 inside t = -log(i), so log x + c = (t + c) + log1p(fma(x, i, -1)) with
 *           log1p(r) = r + r^2 (-1/2 + r / 3) (|r| < 2^-8 + 4e-4); outside that range x is
 *           scaled into binade 0 and e ln2 added in one fma.  Max error of orc_tlogbf 1.05 ulp.
 *
 * Every step is an IEEE-754 single operation or an integer operation on the bit pattern, so any IEEE
 * implementation evaluates them bit-identically.  The tests pin both routines against
 * double-precision exp/log over their whole argument range.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>
#include "orc.h"

static int g_tmode = ORC_TMODE_DET;
static int g_model = ORC_MODEL_SEPARABLE;

void orc_set_modes(int tmode, int model) { g_tmode = tmode; g_model = model; }
void orc_get_modes(int *tmode, int *model) { if (tmode) *tmode = g_tmode; if (model) *model = g_model; }

#define LOG2E_F   1.44269504088896341f
#define LN2_HI_F  0.693145751953125f
#define LN2_LO_F  1.428606765330187e-06f

float orc_qexpf(float x)
{
    if (x != x) return x;
    if (x < -87.0f) return 0.0f;
    if (x > 88.0f) return INFINITY;
    float n = rintf(x * LOG2E_F);
    float r = fmaf(-n, LN2_HI_F, x);
    r = fmaf(-n, LN2_LO_F, r);
    float p = 1.98412698412698413e-04f;          /* 1/7! */
    p = fmaf(p, r, 1.38888888888888889e-03f);    /* 1/6! */
    p = fmaf(p, r, 8.33333333333333333e-03f);    /* 1/5! */
    p = fmaf(p, r, 4.16666666666666667e-02f);    /* 1/4! */
    p = fmaf(p, r, 1.66666666666666667e-01f);    /* 1/3! */
    p = fmaf(p, r, 0.5f);
    p = fmaf(p, r, 1.0f);
    p = fmaf(p, r, 1.0f);
    return ldexpf(p, (int)n);
}

float orc_qlogf(float x)
{
    if (x != x) return x;
    if (x < 0.0f) return NAN;
    if (x == 0.0f) return -INFINITY;
    if (x == INFINITY) return x;
    int esh = 0;
    if (x < 1.17549435082228751e-38f) { x = x * 8388608.0f; esh = -23; }   /* subnormal: scale by 2^23 */
    uint32_t ix;
    memcpy(&ix, &x, sizeof ix);
    ix = ix - 0x3f3504f3u;                   /* 0x3f3504f3 = sqrt(1/2) rounded to float */
    const int e = ((int32_t)ix) >> 23;
    uint32_t mb = (ix & 0x007fffffu) + 0x3f3504f3u;
    float m;
    memcpy(&m, &mb, sizeof m);               /* m in [sqrt(1/2), sqrt(2)) */
    const float f = m - 1.0f;                /* exact */
    float p = -0.07866668701171875f;
    p = fmaf(p, f, 0.12821750342845917f);
    p = fmaf(p, f, -0.131016343832016f);
    p = fmaf(p, f, 0.1418900042772293f);
    p = fmaf(p, f, -0.1662759929895401f);
    p = fmaf(p, f, 0.2000196874141693f);
    p = fmaf(p, f, -0.2500077188014984f);
    p = fmaf(p, f, 0.33333322405815125f);
    p = fmaf(p, f, -0.4999999701976776f);
    const float f2 = f * f;
    const float r = fmaf(p, f2, f);
    const float ef = (float)(e + esh);
    return fmaf(ef, LN2_HI_F, fmaf(ef, LN2_LO_F, r));
}

static const uint32_t TLOG_TAB[64] = {
    0x3fb31609u, 0xbeabf2bcu, 0x3faf4117u, 0xbea0dfd0u, 0x3fab8713u, 0xbe95dea0u, 0x3fa801dcu, 0xbe8b4090u,
    0x3fa4ad84u, 0xbe8100f9u, 0x3fa168b9u, 0xbe6d7a43u, 0x3f9e4786u, 0xbe596d6du, 0x3f9b5080u, 0xbe46104bu,
    0x3f986dc5u, 0xbe32dbf9u, 0x3f95a207u, 0xbe1fe70du, 0x3f92e464u, 0xbe0cf8f1u, 0x3f905c56u, 0xbdf657d0u,
    0x3f8dd751u, 0xbdd24853u, 0x3f8b658bu, 0xbdaeae36u, 0x3f89230du, 0xbd8d356du, 0x3f86cde9u, 0xbd54260eu,
    0x3f84a776u, 0xbd1249f4u, 0x3f82936au, 0xbca3377au, 0x3f800000u, 0x00000000u, 0x3f7a0ec6u, 0x3cc0651fu,
    0x3f72b5b7u, 0x3d5a5c7cu, 0x3f6bb147u, 0x3da94424u, 0x3f652255u, 0x3de30fcdu, 0x3f5eecc3u, 0x3e0da9b7u,
    0x3f59059fu, 0x3e2924a9u, 0x3f5362ddu, 0x3e441625u, 0x3f4e0ad7u, 0x3e5e4e59u, 0x3f49012au, 0x3e77a781u,
    0x3f442f94u, 0x3e88403au, 0x3f3f8f6fu, 0x3e947796u, 0x3f3b3bffu, 0x3ea02977u, 0x3f3712c1u, 0x3eabab45u,
};
#define LN2_F 0.693147182464599609375f

void orc_tlog_table(uint32_t out[64]) { memcpy(out, TLOG_TAB, sizeof TLOG_TAB); }

float orc_tlog_u(float x, float c)
{
    if (x != x || x < 0.0f) return NAN;
    if (x == 0.0f) return -INFINITY;
    if (x == INFINITY) return x;
    int esh = 0;
    if (x < 1.17549435082228751e-38f) { x = x * 8388608.0f; esh = -23; }
    uint32_t bx, mb;
    memcpy(&bx, &x, sizeof bx);
    const uint32_t ix = bx - 0x3f3504f3u;
    mb = bx - (ix & 0xff800000u);
    float m, inv, t;
    memcpy(&m, &mb, sizeof m);                          /* m in [sqrt(1/2), sqrt(2)) */
    const uint32_t j = (ix >> 18) & 31u;
    memcpy(&inv, &TLOG_TAB[2 * j], sizeof inv);
    memcpy(&t, &TLOG_TAB[2 * j + 1], sizeof t);
    const float r = fmaf(m, inv, -1.0f);
    const float r2 = r * r;
    float p = fmaf(r, -0.25f, 0.3333333432674407958984375f);
    p = fmaf(p, r, -0.5f);
    const float lr = fmaf(p, r2, r);
    const float w = (t + lr) + c;
    const float ef = (float)((((int32_t)ix) >> 23) + esh);
    return fmaf(ef, LN2_F, w);
}

float orc_tlogf(float x) { return orc_tlog_u(x, 0.0f); }

/* binade table log of the one-emitter losses: the table is the library's data file
 * (qc-storm_amd/csrc/qcs_logtab.inc, written by logtab_gen.py; its entries are pinned against
 * their definition by the tests) */
static const uint32_t TLOGB_TAB[8192] = {
#include "../qc-storm_amd/csrc/qcs_logtab.inc"
};
#define TLOGB_K   0x3f368000u
#define TLOGB_LO  0x1.6dp-11f   /* 0.712890625 * 2^-10 */
#define TLOGB_HI  0x1.6dp21f    /* 0.712890625 * 2^22 */

void orc_tlogb_table(uint32_t out[8192]) { memcpy(out, TLOGB_TAB, sizeof TLOGB_TAB); }

/* 128 intervals per binade: |r| < 2^-8 + 4e-4 and log1p(r) = r + r^2 (-1/2 + r / 3) */
static float tlogb_core(float x, float c)
{
    uint32_t bx;
    memcpy(&bx, &x, sizeof bx);
    const int g = (((int32_t)(bx - TLOGB_K)) >> 16) + 1280;   /* binade and interval of x */
    float inv, t;
    memcpy(&inv, &TLOGB_TAB[2 * g], sizeof inv);
    memcpy(&t, &TLOGB_TAB[2 * g + 1], sizeof t);
    const float r = fmaf(x, inv, -1.0f);
    const float r2 = r * r;
    const float p = fmaf(r, 0.3333333432674407958984375f, -0.5f);
    const float lr = fmaf(p, r2, r);
    return (t + c) + lr;
}

float orc_tlogb_u(float x, float c)
{
    if (x != x || x < 0.0f) return NAN;
    if (x == 0.0f) return -INFINITY;
    if (x == INFINITY) return x;
    if (x >= TLOGB_LO && x < TLOGB_HI) return tlogb_core(x, c);      /* binades -10..21 */
    int esh = 0;
    if (x < 1.17549435082228751e-38f) { x = x * 8388608.0f; esh = -23; }
    uint32_t bx, mb;
    memcpy(&bx, &x, sizeof bx);
    const uint32_t ix = bx - TLOGB_K;
    mb = bx - (ix & 0xff800000u);                                      /* x 2^-e in binade 0 */
    float xr;
    memcpy(&xr, &mb, sizeof xr);
    return fmaf((float)((((int32_t)ix) >> 23) + esh), LN2_F, tlogb_core(xr, c));
}

float orc_tlogbf(float x) { return orc_tlogb_u(x, 0.0f); }
void  orc_tlogbf_v(const float *x, float *y, long n) { for (long i = 0; i < n; i++) y[i] = orc_tlogbf(x[i]); }
void  orc_tlogf_v(const float *x, float *y, long n) { for (long i = 0; i < n; i++) y[i] = orc_tlogf(x[i]); }

/* The CRLB's division by the step dh = 0.02 as the library's narrow-range form computes it:
 * y = x * 50 corrected by one fma residual (50 is 1 / 0.02f rounded to float).  Test
 * infrastructure: orc_div_dh_scan counts the floats with bit patterns in [b0, b1) where it differs
 * from the IEEE quotient x / 0.02f. */
float orc_div_dh_fast(float x)
{
    const float y = x * 50.0f;
    return fmaf(fmaf(-y, 0.02f, x), 50.0f, y);
}

long long orc_div_dh_scan(uint32_t b0, uint32_t b1)
{
    long long bad = 0;
#pragma omp parallel for reduction(+ : bad) schedule(static)
    for (long long b = b0; b < (long long)b1; b++) {
        const uint32_t u = (uint32_t)b;
        float x, q, r;
        memcpy(&x, &u, sizeof x);
        q = orc_div_dh_fast(x);
        r = x / 0.02f;
        if (memcmp(&q, &r, sizeof q) != 0) bad++;
    }
    return bad;
}

void orc_qexpf_v(const float *x, float *y, long n) { for (long i = 0; i < n; i++) y[i] = orc_qexpf(x[i]); }
void orc_qlogf_v(const float *x, float *y, long n) { for (long i = 0; i < n; i++) y[i] = orc_qlogf(x[i]); }

/* transcendentals used by the FP32 oracle functions */
float orc_exp32(float x) { return g_tmode == ORC_TMODE_DET ? orc_qexpf(x) : expf(x); }
float orc_log32(float x) { return g_tmode == ORC_TMODE_DET ? orc_qlogf(x) : logf(x); }
float orc_log32_u(float x, float c) { return g_tmode == ORC_TMODE_DET ? orc_tlog_u(x, c) : logf(x) + c; }
float orc_log32_ub(float x, float c) { return g_tmode == ORC_TMODE_DET ? orc_tlogb_u(x, c) : logf(x) + c; }
int   orc_model_separable(void) { return g_model == ORC_MODEL_SEPARABLE; }
