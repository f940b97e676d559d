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

void orc_qexpf_v(const float *x, float *y, long n) { for (long i = 0; i < n; i++) y[i] = orc_qexpf(x[i]); }
void orc_qlogf_v(const float *x, float *y, long n) { for (long i = 0; i < n; i++) y[i] = orc_qlogf(x[i]); }

/* transcendentals used by the FP32 oracle functions */
float orc_exp32(float x) { return g_tmode == ORC_TMODE_DET ? orc_qexpf(x) : expf(x); }
float orc_log32(float x) { return g_tmode == ORC_TMODE_DET ? orc_qlogf(x) : logf(x); }
int   orc_model_separable(void) { return g_model == ORC_MODEL_SEPARABLE; }
