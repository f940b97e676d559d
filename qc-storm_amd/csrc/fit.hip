// fit.hip -- BFGS Poisson-MLE Gaussian fitting and CRLB kernels for gfx950.
//
// One ROI per lane of a wave64.  The ROI's photo-electron pixels q and log q live in VGPRs for the
// whole fit; the Gaussian model is evaluated separably (one exp per ROI row and column), so a
// pixel evaluation of the loss costs one FMA for the model, one log and three FMAs for the
// weighted Poisson deviance.  The reference's fixed schedule (gradient, 2 + 10 line-search
// evaluations, gradient, BFGS update; DLL/bfgsMLE_Loc_LD/bfgsMLE_GS2D_core.h:313-426) is run as a
// uniform state machine over its loss evaluations, so the loss body is emitted once and every
// lane of the wave executes the same instruction stream (no divergence: the schedule does not
// depend on the data; the line-search side each lane updates is a per-lane select).
//
// Reference kernels (DLL = MLEbfgs/bfgs_MLE_dll x64/bfgsMLE_dll/bfgsMLE_dll):
//   bfgsMLELoc_Gauss2D   DLL/bfgsMLE_Loc_LD/bfgsMLE_GS2D_core.h:56-310
//   bfgsMLELoc_AS3D      DLL/bfgsMLE_Loc_LD/bfgsMLE_AS3D_core.h:57-337
//   GaussianCRLB_Calc_top DLL/bfgsMLE_Loc_top/LocPrecCalc_GaussianCRLB.h:64-243
#include "qcs_internal.h"
#include "qcs_devmath.h"

namespace qcs {

enum { MODEL_GS2D = 0, MODEL_AS3D = 1, MODEL_2E = 2, MODEL_3E = 3, MODEL_AS3D2E = 4 };

// Parameter layout of each model: amplitude / x / y / sigma_x / sigma_y index of emitter k,
// background index, iteration count, loss clamp, WLE weighting.
template <int MODEL> struct Traits;
template <> struct Traits<MODEL_GS2D> {   // bfgsMLE_GS2D_core.h:25-37
    static constexpr int P = 5, ITERS = 8, K = 1, IB = 4;
    static constexpr bool WEIGHTED = true;
    static constexpr float CLAMP = 0.0f;
    __device__ static constexpr int ia(int) { return 0; }
    __device__ static constexpr int ix(int) { return 1; }
    __device__ static constexpr int iy(int) { return 2; }
    __device__ static constexpr int isx(int) { return 3; }
    __device__ static constexpr int isy(int) { return 3; }
};
template <> struct Traits<MODEL_AS3D> {   // bfgsMLE_AS3D_core.h:25-38, loss clamp :501-505
    static constexpr int P = 6, ITERS = 11, K = 1, IB = 5;
    static constexpr bool WEIGHTED = true;
    static constexpr float CLAMP = 0.01f;
    __device__ static constexpr int ia(int) { return 0; }
    __device__ static constexpr int ix(int) { return 1; }
    __device__ static constexpr int iy(int) { return 2; }
    __device__ static constexpr int isx(int) { return 3; }
    __device__ static constexpr int isy(int) { return 4; }
};
template <> struct Traits<MODEL_2E> {     // bfgsMLE_GS2D_2Emitter_core.h:24-51
    static constexpr int P = 8, ITERS = 13, K = 2, IB = 7;
    static constexpr bool WEIGHTED = true;
    static constexpr float CLAMP = 0.0f;
    __device__ static constexpr int ia(int k) { return 3 * k; }
    __device__ static constexpr int ix(int k) { return 3 * k + 1; }
    __device__ static constexpr int iy(int k) { return 3 * k + 2; }
    __device__ static constexpr int isx(int) { return 6; }
    __device__ static constexpr int isy(int) { return 6; }
};
template <> struct Traits<MODEL_3E> {     // bfgsMLE_GS2D_3Emitter_core.h:24-54 (unweighted)
    static constexpr int P = 11, ITERS = 16, K = 3, IB = 10;
    static constexpr bool WEIGHTED = false;
    static constexpr float CLAMP = 0.0f;
    __device__ static constexpr int ia(int k) { return 3 * k; }
    __device__ static constexpr int ix(int k) { return 3 * k + 1; }
    __device__ static constexpr int iy(int k) { return 3 * k + 2; }
    __device__ static constexpr int isx(int) { return 9; }
    __device__ static constexpr int isy(int) { return 9; }
};
template <> struct Traits<MODEL_AS3D2E> { // bfgsMLE_AS3D_2Emitter_core.h:24-54 (unweighted), clamp :590-594
    static constexpr int P = 11, ITERS = 16, K = 2, IB = 10;
    static constexpr bool WEIGHTED = false;
    static constexpr float CLAMP = 0.01f;
    __device__ static constexpr int ia(int k) { return 5 * k; }
    __device__ static constexpr int ix(int k) { return 5 * k + 1; }
    __device__ static constexpr int iy(int k) { return 5 * k + 2; }
    __device__ static constexpr int isx(int k) { return 5 * k + 3; }
    __device__ static constexpr int isy(int k) { return 5 * k + 4; }
};

FitConst make_fit_const(const qcs_params &p)
{
    FitConst c;
    c.Offset = p.Offset; c.KAdc = p.KAdc; c.QE = p.QE; c.ReadNoise_e = p.ReadNoise_e; c.PixelSize = p.PixelSize;
    c.MultiEmitterFitEn = p.MultiEmitterFitEn; c.WLEEn = p.WLEEn; c.ZDepthCorrFactor = p.ZDepthCorrFactor;
    c.pG[0] = p.p4_XGY; c.pG[1] = p.p3_XGY; c.pG[2] = p.p2_XGY; c.pG[3] = p.p1_XGY; c.pG[4] = p.p0_XGY;
    c.pL[0] = p.p4_XLY; c.pL[1] = p.p3_XLY; c.pL[2] = p.p2_XLY; c.pL[3] = p.p1_XLY; c.pL[4] = p.p0_XLY;
    return c;
}

// ---------------------------------------------------------------------------------------------
// Weighted Poisson deviance, model I = sum_k A_k exp(-(sx_k dx^2 + sy_k dy^2)) + B evaluated
// separably (one exp per emitter and ROI row / column).  Clamps p in place as the reference's
// loss does with its argument array.  The pixel term is w (I - q - q (log I - log q)), formed as
// fma(-q, log I + c, I) with c = 1 - log q precomputed per ROI and the log fused with c
// (logtab_u); the weighted sum runs as sum_r wy[r] sum_c wx[c] t(r, c).
//
// The range check bounds every model value before the pixel loop: with positions below 2^60 in
// magnitude and finite widths each exp term lies in [0, 1], so I lies in [B, sum of the
// amplitudes + B].  Multi-emitter models need B in [2^-126, 2^124] and amplitudes at most 2^124
// (I a positive normal float); the one-emitter models use the binade table log and need B and A
// within [2^-10.5, 2^18.5] so that I stays inside the table.  A lane that fails it (a line-search
// probe gone far out, a background at zero) takes the _any form of the log, whose range and
// special-value handling gives the value the CPU oracle computes.
template <int MODEL, int R, bool FAST, bool SAFE>
__device__ __forceinline__ float pixel_sum(const float (&q)[R * R], const float (&c1)[R * R], const float (&wx)[R],
                                           const float (&wy)[R], const float (&ex)[Traits<MODEL>::K][R],
                                           const float (&ay)[Traits<MODEL>::K][R], float b, const float2 *lt)
{
    using T = Traits<MODEL>;
    constexpr int K = T::K;
    float total = 0.0f;
#pragma unroll
    for (int r = 0; r < R; r++) {
        float acc = 0.0f;
#pragma unroll
        for (int c = 0; c < R; c++) {
            float I;
            if (K == 1) {
                I = __builtin_fmaf(ay[0][r], ex[0][c], b);
            } else {
                I = ay[0][r] * ex[0][c];
#pragma unroll
                for (int k = 1; k < K; k++) I = __builtin_fmaf(ay[k][r], ex[k][c], I);
                I = I + b;
            }
            const float cc = c1[r * R + c];
            float u;
            if (FAST) u = flogf(I) + cc;
            else if (K == 1) u = SAFE ? logtabb_u_any(I, cc, lt) : logtabb_u(I, cc, lt);
            else u = SAFE ? logtab_u_any(I, cc, lt) : logtab_u(I, cc, lt);
            const float t = __builtin_fmaf(-q[r * R + c], u, I);
            acc = T::WEIGHTED ? __builtin_fmaf(wx[c], t, acc) : acc + t;
        }
        total = T::WEIGHTED ? __builtin_fmaf(wy[r], acc, total) : total + acc;
    }
    return total;
}

template <int MODEL, int R, bool FAST>
__device__ __forceinline__ float loss_eval(const float (&q)[R * R], const float (&c1)[R * R], const float (&wx)[R],
                                          const float (&wy)[R], float (&p)[Traits<MODEL>::P], const float2 *lt)
{
    using T = Traits<MODEL>;
    constexpr int P = T::P, K = T::K;
#pragma unroll
    for (int i = 0; i < P; i++) p[i] = (p[i] < T::CLAMP) ? T::CLAMP : p[i];
    const float b = p[T::IB] * 64.0f;
    // one-emitter models: every model value inside the binade table, I in [b, a + b]
    constexpr float BLO = (K == 1) ? kLogTabBLo : 1.17549435082228751e-38f;
    constexpr float AHI = (K == 1) ? kLogTabBHi * 0.125f : 0x1p124f;
    bool inrange = (b >= BLO) && (b <= AHI);
    float ex[K][R], ay[K][R];
#pragma unroll
    for (int k = 0; k < K; k++) {
        const float a = p[T::ia(k)] * 128.0f;
        const float sx = p[T::isx(k)] * 0.1f;
        const float sy = p[T::isy(k)] * 0.1f;
        inrange = inrange && (a <= AHI) && (sx <= 0x1p124f) && (sy <= 0x1p124f) &&
                  (__builtin_fabsf(p[T::ix(k)]) <= 0x1p60f) && (__builtin_fabsf(p[T::iy(k)]) <= 0x1p60f);
#pragma unroll
        for (int c = 0; c < R; c++) {
            const float dx = ((float)c + 0.5f) - p[T::ix(k)];
            ex[k][c] = vexp_nonpos<FAST>((-(dx * dx)) * sx);
        }
#pragma unroll
        for (int r = 0; r < R; r++) {
            const float dy = ((float)r + 0.5f) - p[T::iy(k)];
            ay[k][r] = a * vexp_nonpos<FAST>((-(dy * dy)) * sy);
        }
    }
    if (FAST || inrange) return pixel_sum<MODEL, R, FAST, false>(q, c1, wx, wy, ex, ay, b, lt);
    return pixel_sum<MODEL, R, FAST, true>(q, c1, wx, wy, ex, ay, b, lt);
}

// ConstructD0 (bfgsMLE_core.h:87-182).  D is the lane's inverse-Hessian approximation, stored in
// LDS with element k at D[k * 64] (lane-interleaved: conflict-free, off the VGPR budget).
template <int P>
__device__ __forceinline__ void bfgs_update(float *D, const float (&s)[P], const float (&y)[P])
{
    float sy = 0.0f;
#pragma unroll
    for (int k = 0; k < P; k++) sy = __builtin_fmaf(s[k], y[k], sy);
    sy = 1.0f / sy;
    float ty[P], ts[P];
#pragma unroll
    for (int k = 0; k < P; k++) { ty[k] = y[k] * sy; ts[k] = s[k] * sy; }
    float A[P * P], T[P * P];
#pragma unroll
    for (int r = 0; r < P; r++)
#pragma unroll
        for (int c = 0; c < P; c++) A[r * P + c] = __builtin_fmaf(-s[r], ty[c], (r == c) ? 1.0f : 0.0f);
#pragma unroll
    for (int c = 0; c < P; c++) {
        float Dc[P];
#pragma unroll
        for (int k = 0; k < P; k++) Dc[k] = D[(k * P + c) * 64];
#pragma unroll
        for (int r = 0; r < P; r++) {
            float acc = 0.0f;
#pragma unroll
            for (int k = 0; k < P; k++) acc = __builtin_fmaf(A[r * P + k], Dc[k], acc);
            T[r * P + c] = acc;
        }
    }
#pragma unroll
    for (int r = 0; r < P; r++)
#pragma unroll
        for (int c = 0; c < P; c++) {
            float acc = 0.0f;
#pragma unroll
            for (int k = 0; k < P; k++) acc = __builtin_fmaf(T[r * P + k], A[c * P + k], acc);
            D[(r * P + c) * 64] = acc + s[r] * ts[c];
        }
}

template <int P>
__device__ __forceinline__ void normalize_dir(float (&d)[P])
{
    float tot = 0.0f;
#pragma unroll
    for (int k = 0; k < P; k++) tot = tot + __builtin_fabsf(d[k]);
    tot = (float)P / tot;
#pragma unroll
    for (int k = 0; k < P; k++) d[k] = d[k] * tot;
}

enum { PH_G0 = 0, PH_G1 = 1, PH_L0 = 2, PH_L1 = 3, PH_BS = 4 };

// Fixed-schedule BFGS (MLELocalization_GS2D, bfgsMLE_GS2D_core.h:313-426, with the gradient of
// NumericalGradient_GS2D :429-454).  x holds the initial vector on entry, the result on exit.
// Schedule: gradient (2P evaluations); per iteration: probes at 1e-4 and 1, ten bisection
// evaluations (eleven halvings of [1e-5, 1]), step, clamp >= 0.01, and -- except after the
// last iteration -- gradient, BFGS update and new direction.
template <int MODEL, int R, bool FAST>
__device__ __forceinline__ void bfgs_fit(const float (&q)[R * R], const float (&c1)[R * R], const float (&wx)[R],
                                            const float (&wy)[R], float (&x)[Traits<MODEL>::P], float *cold,
                                            const float2 *lt)
{
    constexpr int P = Traits<MODEL>::P, IT = Traits<MODEL>::ITERS;
    constexpr int NEVAL = 2 * P + IT * 12 + (IT - 1) * 2 * P;
    // cold state in LDS (element k at cold[k * 64]): D (P*P), previous gradient (P), step (P)
    float *D = cold, *gprev_l = cold + P * P * 64, *s_l = cold + (P * P + P) * 64;
    float g[P], d[P], t[P];
#pragma unroll
    for (int r = 0; r < P; r++)
#pragma unroll
        for (int c = 0; c < P; c++) D[(r * P + c) * 64] = (r == c) ? 1.0f : 0.0f;
#pragma unroll
    for (int k = 0; k < P; k++) { t[k] = x[k]; g[k] = 0.0f; d[k] = 0.0f; }
    float L0 = 0.0f, L1 = 0.0f, lo = 0.0f, hi = 0.0f, ln = 0.0f;
    bool sel = false;
    int phase = PH_G0, gi = 0, bsn = 0;
    bool first = true;

    for (int e = 0; e < NEVAL; e++) {
        float pr[P];
        if (phase <= PH_G1) {
            const float delta = (phase == PH_G0) ? -0.002f : 0.004f;
#pragma unroll
            for (int k = 0; k < P; k++) {
                if (k == gi) t[k] = t[k] + delta;
                pr[k] = t[k];
            }
        } else {
            float step = (phase == PH_L0) ? 0.0001f : 1.0f;
            if (phase == PH_BS) step = sel ? hi : lo;
#pragma unroll
            for (int k = 0; k < P; k++) pr[k] = __builtin_fmaf(d[k], step, x[k]);
        }

        const float L = loss_eval<MODEL, R, FAST>(q, c1, wx, wy, pr, lt);

        if (phase == PH_G0) {
#pragma unroll
            for (int k = 0; k < P; k++) t[k] = pr[k];
            ln = L;
            phase = PH_G1;
        } else if (phase == PH_G1) {
#pragma unroll
            for (int k = 0; k < P; k++) {
                t[k] = pr[k];
                if (k == gi) {
                    g[k] = (L - ln) * 250.0f;
                    t[k] = t[k] - 0.002f;
                }
            }
            gi++;
            phase = PH_G0;
            if (gi == P) {
                if (first) {
                    first = false;
#pragma unroll
                    for (int k = 0; k < P; k++) d[k] = -g[k];
                } else {
                    float y[P], s[P];
#pragma unroll
                    for (int k = 0; k < P; k++) { y[k] = g[k] - gprev_l[k * 64]; s[k] = s_l[k * 64]; }
                    bfgs_update<P>(D, s, y);
#pragma unroll
                    for (int r = 0; r < P; r++) {     // MatMulVector (bfgsMLE_core.h:184-204)
                        float acc = 0.0f;
#pragma unroll
                        for (int c = 0; c < P; c++) acc = __builtin_fmaf(-D[(r * P + c) * 64], g[c], acc);
                        d[r] = acc;
                    }
                }
                normalize_dir<P>(d);
                lo = 0.00001f;
                hi = 1.0f;
                phase = PH_L0;
            }
        } else if (phase == PH_L0) {
            L0 = L;
            phase = PH_L1;
        } else {
            if (phase == PH_L1) {
                L1 = L;
                bsn = 0;
            } else {
                if (sel) L1 = L; else L0 = L;
                bsn++;
            }
            sel = (L0 < L1);
            const float mid = (lo + hi) * 0.5f;
            if (sel) hi = mid; else lo = mid;
            phase = PH_BS;
            if (bsn == 10) {
                const float scale = (lo + hi) * 0.5f;
#pragma unroll
                for (int k = 0; k < P; k++) {
                    const float sk = d[k] * scale;
                    s_l[k * 64] = sk;
                    x[k] = x[k] + sk;
                }
#pragma unroll
                for (int k = 0; k < P; k++) {
                    x[k] = (x[k] < 0.01f) ? 0.01f : x[k];   // IninfConstrain (bfgsMLE_core.h:76-85)
                    gprev_l[k * 64] = g[k];
                    t[k] = x[k];
                }
                gi = 0;
                phase = PH_G0;
            }
        }
    }
}

// ---------------------------------------------------------------------------------------------
// Per-ROI setup shared by the fitting kernels: photo-electrons q, the loss's data term
// c1 = 1 - log q, header.
template <int R>
__device__ __forceinline__ void load_roi(const uint16_t *__restrict__ rec, const FitConst &fc, float (&q)[R * R],
                                         float (&c1)[R * R], int &xo, int &yo, int &fr)
{
    // the record is R*(R+1) 16-bit words; read it as 32-bit words (R*(R+1) is even)
    constexpr int NW = R * (R + 1) / 2;
    const uint32_t *w = reinterpret_cast<const uint32_t *>(rec);
    uint32_t v[NW];
#pragma unroll
    for (int k = 0; k < NW; k++) v[k] = __builtin_nontemporal_load(w + k);
#pragma unroll
    for (int i = 0; i < R * R; i++) {
        const uint32_t word = v[i >> 1];
        const float adu = (float)((i & 1) ? (word >> 16) : (word & 0xffffu));
        const float e = (adu - fc.Offset) * fc.KAdc;     // bfgsMLE_GS2D_core.h:144
        q[i] = (e > 1.0f) ? e : 1.0f;
        c1[i] = 1.0f - qlogf_normal(q[i]);   // q >= 1
    }
    auto word_at = [&](int idx) -> int {
        const uint32_t word = v[idx >> 1];
        return (int)((idx & 1) ? (word >> 16) : (word & 0xffffu));
    };
    xo = word_at(R * R + 0);
    yo = word_at(R * R + 1);
    fr = word_at(R * R + 2) + 65536 * word_at(R * R + 3);
}

// WLEWeightsCalc (bfgsMLE_core.h:206-238), separable form: w(r, c) = wy[r] * wx[c]
template <int R>
__device__ __forceinline__ void wle_weights(float sgx, float sgy, bool mlefit, float (&wx)[R], float (&wy)[R])
{
    const float ctr = (float)R / 2.0f;
    const float sx1 = 1.0f / ((2.0f * sgx) * sgx);
    const float sy1 = 1.0f / ((2.0f * sgy) * sgy);
#pragma unroll
    for (int c = 0; c < R; c++) {
        const float d = ((float)c + 0.5f) - ctr;
        wx[c] = mlefit ? 1.0f : qexpf(-((d * d) * sx1));
        wy[c] = mlefit ? 1.0f : qexpf(-((d * d) * sy1));
    }
}

// Border sums of the ROI (PreEstimation_GS2D, bfgsMLE_GS2D_core.h:529-546).
template <int R>
__device__ __forceinline__ float border_bkg(const float (&q)[R * R])
{
    float m1 = 0.0f, m2 = 0.0f, m3 = 0.0f, m4 = 0.0f;
#pragma unroll
    for (int k = 0; k < R; k++) {
        m1 = m1 + q[k * R + 0];
        m2 = m2 + q[k * R + R - 1];
        m3 = m3 + q[0 * R + k];
        m4 = m4 + q[(R - 1) * R + k];
    }
    m1 = m1 < m2 ? m1 : m2;
    m3 = m3 < m4 ? m3 : m4;
    return ((m1 + m3) / 2.0f) / (float)R;
}

__device__ __forceinline__ float zpoly(const float (&pc)[5], float d, float corr)
{
    // bfgsMLE_AS3D_core.h:266-279 with d^k formed by multiplication
    const float d2 = d * d, d3 = d2 * d, d4 = d2 * d2;
    float z = pc[0] * d4;
    z = __builtin_fmaf(pc[1], d3, z);
    z = __builtin_fmaf(pc[2], d2, z);
    z = __builtin_fmaf(pc[3], d, z);
    z = z + pc[4];
    return z * corr;
}

// One-emitter fitting kernel.  Lane i fits ROI idx[i] (or i when idx is null) and writes its
// 12-float row; flag[roi] (if non-null) is set to 1 when a rejected fit qualifies for the
// two-emitter stage (valid background and SNR > 6, bfgsMLE_GS2D_core.h:279-297).
// Up to 9x9 the register budget is held to two waves per SIMD (256 VGPRs).  The 7x7 exact-math
// kernel fits without spills; at 9x9 the kernel spills 30 (2D) or 65 (elliptical) registers outside
// its loss loop and still runs 1.5-1.6x faster than at its natural one wave.
// Waves per work-group: they share the 32 KB log table in LDS.  With each wave's BFGS state
// (9 KB for the 2D model, 12 KB elliptical) a CU holds two groups of four 2D waves, or one group
// of eight elliptical ones, i.e. two waves per SIMD up to 9x9.
template <int MODEL, int R> constexpr int fit_wg_waves() { return (MODEL == MODEL_AS3D && R <= 9) ? 8 : 4; }

template <int MODEL, int R, bool FAST>
__global__ void __launch_bounds__((64 * fit_wg_waves<MODEL, R>()))
__attribute__((amdgpu_waves_per_eu(R <= 9 ? 2 : 1, R <= 9 ? 2 : 10)))
fit_1e_kernel(const uint16_t *__restrict__ rois, const float *__restrict__ wle,
                                                    const int *__restrict__ idx, int n, const int *__restrict__ n_dev,
                                                    float *__restrict__ loc, int *__restrict__ flag, FitConst fc)
{
    constexpr int P = Traits<MODEL>::P;
    constexpr int HALF = R / 2;
    __shared__ float2 ltab[kLogTabBN];
    log_tabb_load(ltab);
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (n_dev) n = *n_dev;
    if (i >= n) return;
    const int roi = idx ? idx[i] : i;

    float q[R * R], c1[R * R];
    int xo, yo, fr;
    load_roi<R>(rois + (size_t)roi * (R * (R + 1)), fc, q, c1, xo, yo, fr);

    const float *wp = wle + (size_t)roi * QCS_WLE_FLOATS;
    const float sl = wp[1], sr = wp[2], su = wp[3], sd = wp[4];
    const int mt = (int)wp[5];
    const bool mlefit = (mt == QCS_MOLTYPE_MLEFIT);

    const float bkg = border_bkg<R>(q);
    const float amp = q[HALF * R + HALF] - bkg;
    float x[P], wx[R], wy[R];
    if (MODEL == MODEL_GS2D) {
        // bfgsMLELoc_Gauss2D :188-228, PreEstimation_GS2D :548-574
        float ws = fminf(fminf(sl, sr), fminf(su, sd));
        ws = fminf(ws, ((float)R / 2.0f) / 2.35f);
        const float ps = 1.0f / ((2.0f * ws) * ws);
        x[0] = amp * 0.0078125f;
        x[1] = (float)HALF + 0.5f;
        x[2] = (float)HALF + 0.5f;
        x[3] = ps * 10.0f;
        x[4] = bkg * 0.015625f;
        if (!mlefit) {
            ws = ws / 1.2f;
            ws = fminf(ws, ((float)R / 2.0f) / 2.35f);
        }
        if (fc.WLEEn == 0 && mt != QCS_MOLTYPE_MULTIFIT) ws = (float)R;
        wle_weights<R>(ws, ws, mlefit, wx, wy);
    } else {
        // bfgsMLELoc_AS3D :188-228, PreEstimation_AS3D :578-606
        float wsx = fminf(sl, sr), wsy = fminf(su, sd);
        x[0] = amp * 0.0078125f;
        x[1] = (float)HALF + 0.5f;
        x[2] = (float)HALF + 0.5f;
        x[3] = 10.0f * (0.5f / (wsx * wsx));
        x[4] = 10.0f * (0.5f / (wsy * wsy));
        x[5] = bkg * 0.015625f;
        if (!mlefit) {
            const float cap = ((float)R / 1.6f) / 2.35f;
            wsx = fminf(wsx, cap);
            wsy = fminf(wsy, cap);
        }
        if (fc.WLEEn == 0 && mt != QCS_MOLTYPE_MULTIFIT) { wsx = (float)R; wsy = (float)R; }
        wle_weights<R>(wsx, wsy, mlefit, wx, wy);
    }

    __shared__ float cold[fit_wg_waves<MODEL, R>()][(P * P + 2 * P) * 64];   // per wave
    bfgs_fit<MODEL, R, FAST>(q, c1, wx, wy, x, &cold[threadIdx.x >> 6][threadIdx.x & 63], ltab);

    float row[QCS_REC_FLOATS];
    bool ok;
    float snr;
    bool valid;
    const float QE = fc.QE;
    if (MODEL == MODEL_GS2D) {
        // bfgsMLELoc_Gauss2D :236-308
        const float lb = (float)R / 2.0f - 1.5f, ub = (float)R / 2.0f + 1.5f;
        const float smax = fc.MultiEmitterFitEn ? ((float)R / 2.0f) / 2.35f : ((float)R / 1.5f) / 2.35f;
        const float pp = (x[0] * 128.0f) / QE;
        const float bg = (x[4] * 64.0f) / QE;
        float fx = x[1], fy = x[2];
        float sg = sqrtf(0.5f / (x[3] * 0.1f));
        fx = (fx != fx) ? -1.0f : fx;
        fy = (fy != fy) ? -1.0f : fy;
        sg = (sg != sg) ? -1.0f : sg;
        snr = (sqrtf(QE) * pp) / sqrtf(pp + bg);
        const float tp = (((pp * 2.0f) * 3.141593f) * sg) * sg;
        valid = !(bg != bg);
        ok = valid && snr > 4.0f && sg <= smax && sg >= 0.6f && fx > lb && fx < ub && fy > lb && fy < ub;
        row[0] = pp; row[1] = (fx - (float)HALF) + (float)xo; row[2] = (fy - (float)HALF) + (float)yo;
        row[3] = 0.0f; row[4] = sg; row[5] = sg; row[6] = tp; row[7] = bg; row[8] = snr;
        row[9] = 0.0f; row[10] = 0.0f; row[11] = (float)fr;
    } else {
        // bfgsMLELoc_AS3D :237-335
        const float lb = (float)R / 2.0f - 2.5f, ub = (float)R / 2.0f + 2.5f;
        const float smax = fc.MultiEmitterFitEn ? ((float)R / 2.0f) / 2.35f : ((float)R / 1.5f) / 2.35f;
        const float pe = x[0] * 128.0f, be = x[5] * 64.0f;
        float fx = x[1], fy = x[2];
        float sx = sqrtf(0.5f / (x[3] * 0.1f));
        float sy = sqrtf(0.5f / (x[4] * 0.1f));
        fx = (fx != fx) ? -1.0f : fx;
        fy = (fy != fy) ? -1.0f : fy;
        sx = (sx != sx) ? -1.0f : sx;
        sy = (sy != sy) ? -1.0f : sy;
        snr = pe / sqrtf(pe + be);
        const float pp = pe / QE, bg = be / QE;
        const float tp = (((pp * 2.0f) * 3.141593f) * sx) * sy;
        const float dif = (sx * sx) - (sy * sy);
        const float z = (dif >= 0.0f) ? zpoly(fc.pG, dif, fc.ZDepthCorrFactor) : zpoly(fc.pL, dif, fc.ZDepthCorrFactor);
        const float ms = (sx + sy) / 2.0f;
        valid = !(bg != bg);
        ok = valid && snr > 3.5f && ms > 0.8f && ms < smax && fx > lb && fx < ub && fy > lb && fy < ub;
        row[0] = pp; row[1] = (fx - (float)HALF) + (float)xo; row[2] = (fy - (float)HALF) + (float)yo;
        row[3] = z; row[4] = sx; row[5] = sy; row[6] = tp; row[7] = bg; row[8] = snr;
        row[9] = 0.0f; row[10] = 0.0f; row[11] = (float)fr;
    }
    if (!ok) {
#pragma unroll
        for (int k = 0; k < QCS_REC_FLOATS; k++) row[k] = 0.0f;
    }
    float4 *dst = reinterpret_cast<float4 *>(loc + (size_t)roi * QCS_REC_FLOATS);
    dst[0] = make_float4(row[0], row[1], row[2], row[3]);
    dst[1] = make_float4(row[4], row[5], row[6], row[7]);
    dst[2] = make_float4(row[8], row[9], row[10], row[11]);
    if (flag) flag[roi] = (!ok && valid && snr > 6.0f) ? 1 : 0;
}

#include "fit_multi.inc"

// ---------------------------------------------------------------------------------------------
// Diagonal-Fisher CRLB of x and y (GaussianCRLB_Calc_top, LocPrecCalc_GaussianCRLB.h:64-128):
// numeric second derivatives of the model and its log with dh = 0.02 on a G x G grid.
template <int G, bool NARROW>
__device__ __forceinline__ float crlb_psf(int r, int c, float peak, float ix, float iy, float x0, float y0, float B)
{
    const float cy = (float)(r - G / 2), cx = (float)(c - G / 2);
    const float dx = cx - x0, dy = cy - y0;
    const float arg = -__builtin_fmaf(dx * dx, ix, (dy * dy) * iy);
    return __builtin_fmaf(peak, NARROW ? qexpf_nonpos(arg) : qexpf(arg), B);
}

// x / 0.02f.  NARROW: as y = x * 50 corrected by one fma residual, which equals the IEEE quotient
// for x = 0 and every |x| in [1e-30, 1e30] (checked over all floats); the Fisher sums' numerators
// stay there when every model value lies in [1, 1e20].
template <bool NARROW>
__device__ __forceinline__ float div_dh(float x)
{
    if (!NARROW) return x / 0.02f;
    const float y = x * 50.0f;
    return __builtin_fmaf(__builtin_fmaf(-y, 0.02f, x), 50.0f, y);
}

// Diagonal Fisher sums of x and y.  NARROW (every model value in [1, 1e20], the argument of each
// exponential <= 0): qexpf_nonpos, qlogf_normal and div_dh<true> give the same bits as the
// general forms there.
template <int G, bool NARROW>
__device__ __forceinline__ void crlb_fisher(float peak, float ix, float iy, float be, float &fx, float &fy)
{
    const float dh = 0.02f;
    const float xp = 0.0f + dh, xn = (0.0f + dh) - 2.0f * dh;
    auto lg = [](float m) { return NARROW ? qlogf_normal(m) : qlogf(m); };
    fx = 0.0f;
    fy = 0.0f;
    for (int r = 0; r < G; r++) {
#pragma unroll
        for (int c = 0; c < G; c++) {
            const float m0 = crlb_psf<G, NARROW>(r, c, peak, ix, iy, 0.0f, 0.0f, be);
            const float l0 = lg(m0);
            {
                const float mp = crlb_psf<G, NARROW>(r, c, peak, ix, iy, xp, 0.0f, be);
                const float mn = crlb_psf<G, NARROW>(r, c, peak, ix, iy, xn, 0.0f, be);
                const float d2 = div_dh<NARROW>(div_dh<NARROW>(__builtin_fmaf(-2.0f, m0, mp) + mn));
                const float d2l = div_dh<NARROW>(div_dh<NARROW>(__builtin_fmaf(-2.0f, l0, lg(mp)) + lg(mn)));
                fx = fx + __builtin_fmaf(-m0, d2l, d2);
            }
            {
                const float mp = crlb_psf<G, NARROW>(r, c, peak, ix, iy, 0.0f, xp, be);
                const float mn = crlb_psf<G, NARROW>(r, c, peak, ix, iy, 0.0f, xn, be);
                const float d2 = div_dh<NARROW>(div_dh<NARROW>(__builtin_fmaf(-2.0f, m0, mp) + mn));
                const float d2l = div_dh<NARROW>(div_dh<NARROW>(__builtin_fmaf(-2.0f, l0, lg(mp)) + lg(mn)));
                fy = fy + __builtin_fmaf(-m0, d2l, d2);
            }
        }
    }
}

template <int G>
__global__ void __launch_bounds__(64) crlb_kernel(float *__restrict__ loc, int n, const int *__restrict__ n_dev, float QE,
                                                  float RN, float PS)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (n_dev) n = *n_dev;
    if (i >= n) return;
    float *row = loc + (size_t)i * QCS_REC_FLOATS;
    float te = row[6] * QE, be = row[7] * QE;
    const float sx = row[4], sy = row[5];
    float stdx = 100.0f, stdy = 100.0f;
    if (te > 1.0f && be > 1.0f && (double)sx > 0.1 && (double)sy > 0.1 && sx < 10.0f && sy < 10.0f) {
        be = __builtin_fmaf(be, QE, RN * RN) / QE;
        const float pi = 3.141592654f;
        const float peak = (1.0f / (((2.0f * pi) * sx) * sy)) * te;
        const float ix = 1.0f / ((2.0f * sx) * sx);
        const float iy = 1.0f / ((2.0f * sy) * sy);
        float fx, fy;
        // model values lie in [be, peak + be] with be > 1
        if (peak + be < 1e20f) crlb_fisher<G, true>(peak, ix, iy, be, fx, fy);
        else crlb_fisher<G, false>(peak, ix, iy, be, fx, fy);
        fx = (fx == 0.0f) ? 1.0f : fx;
        fy = (fy == 0.0f) ? 1.0f : fy;
        stdx = sqrtf(1.0f / fx) * PS;
        stdy = sqrtf(1.0f / fy) * PS;
    }
    row[9] = stdx;
    row[10] = stdy;
}

// ---------------------------------------------------------------------------------------------
template <int MODEL, int R, bool FAST>
static void launch_1e(hipStream_t s, const uint16_t *rois, const float *wle, const int *idx, int n, const int *n_dev,
                      float *loc, int *flag, const FitConst &fc)
{
    constexpr int T = 64 * fit_wg_waves<MODEL, R>();
    hipLaunchKernelGGL((fit_1e_kernel<MODEL, R, FAST>), dim3((unsigned)((n + T - 1) / T)), dim3(T), 0, s, rois, wle, idx, n,
                       n_dev, loc, flag, fc);
}

template <int MODEL, bool FAST>
static int launch_1e_r(int R, hipStream_t s, const uint16_t *rois, const float *wle, const int *idx, int n,
                       const int *n_dev, float *loc, int *flag, const FitConst &fc)
{
    switch (R) {
    case 5: launch_1e<MODEL, 5, FAST>(s, rois, wle, idx, n, n_dev, loc, flag, fc); break;
    case 7: launch_1e<MODEL, 7, FAST>(s, rois, wle, idx, n, n_dev, loc, flag, fc); break;
    case 9: launch_1e<MODEL, 9, FAST>(s, rois, wle, idx, n, n_dev, loc, flag, fc); break;
    case 11: launch_1e<MODEL, 11, FAST>(s, rois, wle, idx, n, n_dev, loc, flag, fc); break;
    case 13: launch_1e<MODEL, 13, FAST>(s, rois, wle, idx, n, n_dev, loc, flag, fc); break;
    case 15: launch_1e<MODEL, 15, FAST>(s, rois, wle, idx, n, n_dev, loc, flag, fc); break;
    case 17: launch_1e<MODEL, 17, FAST>(s, rois, wle, idx, n, n_dev, loc, flag, fc); break;
    default: return fail(QCS_ERR_ARG, "ROISize %d not supported by the fitter (odd, 5..17)", R);
    }
    QCS_HIP(hipGetLastError());
    return QCS_OK;
}

int fit_launch_1e(const qcs_params &p, const uint16_t *d_rois, const float *d_wle, const int *d_idx, int n,
                  const int *d_n, float *d_loc, int *d_flag, hipStream_t s)
{
    if (n <= 0) return QCS_OK;
    const FitConst fc = make_fit_const(p);
    const bool as3d = (p.LocType == QCS_LOCTYPE_AS3D);
    const bool fast = (p.FitMath == QCS_MATH_FAST);
    if (as3d)
        return fast ? launch_1e_r<MODEL_AS3D, true>(p.ROISize, s, d_rois, d_wle, d_idx, n, d_n, d_loc, d_flag, fc)
                    : launch_1e_r<MODEL_AS3D, false>(p.ROISize, s, d_rois, d_wle, d_idx, n, d_n, d_loc, d_flag, fc);
    return fast ? launch_1e_r<MODEL_GS2D, true>(p.ROISize, s, d_rois, d_wle, d_idx, n, d_n, d_loc, d_flag, fc)
                : launch_1e_r<MODEL_GS2D, false>(p.ROISize, s, d_rois, d_wle, d_idx, n, d_n, d_loc, d_flag, fc);
}

template <int MODEL, bool FAST>
static int launch_me_r(int R, dim3 grid, hipStream_t s, const uint16_t *rois, const float *wle, const int *list,
                       const int *n_dev, float *loc, float *second, int *fs, int *ff, const FitConst &fc)
{
    switch (R) {
    case 5: hipLaunchKernelGGL((fit_me_kernel<MODEL, 5, FAST>), grid, dim3(64), 0, s, rois, wle, list, n_dev, loc, second, fs, ff, fc); break;
    case 7: hipLaunchKernelGGL((fit_me_kernel<MODEL, 7, FAST>), grid, dim3(64), 0, s, rois, wle, list, n_dev, loc, second, fs, ff, fc); break;
    case 9: hipLaunchKernelGGL((fit_me_kernel<MODEL, 9, FAST>), grid, dim3(64), 0, s, rois, wle, list, n_dev, loc, second, fs, ff, fc); break;
    case 11: hipLaunchKernelGGL((fit_me_kernel<MODEL, 11, FAST>), grid, dim3(64), 0, s, rois, wle, list, n_dev, loc, second, fs, ff, fc); break;
    case 13: hipLaunchKernelGGL((fit_me_kernel<MODEL, 13, FAST>), grid, dim3(64), 0, s, rois, wle, list, n_dev, loc, second, fs, ff, fc); break;
    default: return fail(QCS_ERR_ARG, "ROISize %d not supported by the multi-emitter fitter (odd, 5..13)", R);
    }
    QCS_HIP(hipGetLastError());
    return QCS_OK;
}

// multi-emitter stage over list[j], j < *n_dev (grid sized for n_max)
int fit_launch_me(const qcs_params &p, int model, const uint16_t *d_rois, const float *d_wle, const int *d_list, int n_max,
                  const int *d_n, float *d_loc, float *d_second, int *d_fs, int *d_ff, hipStream_t s)
{
    if (n_max <= 0) return QCS_OK;
    const FitConst fc = make_fit_const(p);
    const dim3 grid((n_max + 63) / 64);
    const bool fast = (p.FitMath == QCS_MATH_FAST);
    const int R = p.ROISize;
    switch (model) {
    case MODEL_2E:
        return fast ? launch_me_r<MODEL_2E, true>(R, grid, s, d_rois, d_wle, d_list, d_n, d_loc, d_second, d_fs, d_ff, fc)
                    : launch_me_r<MODEL_2E, false>(R, grid, s, d_rois, d_wle, d_list, d_n, d_loc, d_second, d_fs, d_ff, fc);
    case MODEL_3E:
        return fast ? launch_me_r<MODEL_3E, true>(R, grid, s, d_rois, d_wle, d_list, d_n, d_loc, d_second, d_fs, d_ff, fc)
                    : launch_me_r<MODEL_3E, false>(R, grid, s, d_rois, d_wle, d_list, d_n, d_loc, d_second, d_fs, d_ff, fc);
    default:
        return fast ? launch_me_r<MODEL_AS3D2E, true>(R, grid, s, d_rois, d_wle, d_list, d_n, d_loc, d_second, d_fs, d_ff, fc)
                    : launch_me_r<MODEL_AS3D2E, false>(R, grid, s, d_rois, d_wle, d_list, d_n, d_loc, d_second, d_fs, d_ff, fc);
    }
}

int crlb_launch(const qcs_params &p, float *d_loc, int n, const int *d_n, hipStream_t s)
{
    if (n <= 0) return QCS_OK;
    const dim3 grid((n + 63) / 64);
    if (p.LocType == QCS_LOCTYPE_AS3D)
        hipLaunchKernelGGL(crlb_kernel<15>, grid, dim3(64), 0, s, d_loc, n, d_n, p.QE, p.ReadNoise_e, p.PixelSize);
    else
        hipLaunchKernelGGL(crlb_kernel<9>, grid, dim3(64), 0, s, d_loc, n, d_n, p.QE, p.ReadNoise_e, p.PixelSize);
    QCS_HIP(hipGetLastError());
    return QCS_OK;
}

// ---------------------------------------------------------------------------------------------
// Multi-emitter cascade (LDLocData_TypeDef::BFGS_MLELocalization, bfgs_top.cu:26-204, with
// gpuMoleculeFitClasify, bfgs_top_MoleculeClasify.cu:26-92).  The reference builds its lists
// with atomic counters (arbitrary order); here every list is a stable compaction (exclusive
// scan of flags), so list order and the rows appended for second emitters are deterministic:
//   1E list  = ROIs with WLE type <= WLEFit, in index order
//   2E list  = ROIs with WLE type MultiFit, in index order, then rejected 1E fits with SNR > 6
//   3E list  = rejected first emitters of the 2E stage, in 2E-list order (2D only)
//   appended = accepted second emitters of the 2E stage, then of the 3E stage, in list order
// All counts stay on the device; one read-back at the end.

__global__ void cascade_classify_kernel(const float *__restrict__ wle, int n, int *__restrict__ f1, int *__restrict__ f2)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int mt = (int)wle[(size_t)i * QCS_WLE_FLOATS + QCS_WLE_FLUOTYPE];
    f1[i] = (mt <= QCS_MOLTYPE_WLEFIT) ? 1 : 0;
    f2[i] = (mt <= QCS_MOLTYPE_WLEFIT) ? 0 : 1;
}

// cnt layout: [0] n1, [1] n2 classified, [2] n2 total, [3] n3, [4] seconds from 2E, [5] seconds
// from 3E, [6] rejected, [7] rows, [8] loc_cap
__global__ void cascade_count_kernel(const int *__restrict__ tile_sums, int ntiles, int *__restrict__ cnt, int slot, int mode)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const int total = tile_sums[ntiles];
    if (mode == 0) cnt[slot] = total;
    else if (mode == 1) cnt[slot] += total;   // accumulate (2E list = classified + re-routed)
}

__global__ void cascade_rows_kernel(int n, int *__restrict__ cnt)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    int extra = cnt[4] + cnt[5];
    const int room = cnt[8] - n;
    extra = extra < room ? extra : room;
    cnt[7] = n + extra;
}

// append second-emitter rows: row = n + base + offs[j] for flagged j (bounded by the capacity)
__global__ void __launch_bounds__(256) append_kernel(const int *__restrict__ flags, const int *__restrict__ offs,
                                                     const int *__restrict__ n_list, const float *__restrict__ second,
                                                     const int *__restrict__ cnt, int base_slot, int n, int loc_cap,
                                                     float *__restrict__ loc)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= *n_list || !flags[j]) return;
    const int base = base_slot >= 0 ? cnt[base_slot] : 0;
    const int row = n + base + offs[j];
    if (row >= loc_cap) return;
    const float4 *src = reinterpret_cast<const float4 *>(second + (size_t)j * QCS_REC_FLOATS);
    float4 *dst = reinterpret_cast<float4 *>(loc + (size_t)row * QCS_REC_FLOATS);
    dst[0] = src[0];
    dst[1] = src[1];
    dst[2] = src[2];
}

__global__ void cascade_set_kernel(int *p, int v) { *p = v; }

// flags of entries j >= *n_list are cleared so a scan over n entries counts only the list
__global__ void clear_tail_kernel(int *__restrict__ f, int n, const int *__restrict__ n_list)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n && j >= *n_list) f[j] = 0;
}

size_t cascade_scratch_bytes(int n)
{
    auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const size_t ni = (size_t)(n > 0 ? n : 1);
    return up(64) + 3 * up(sizeof(int) * ni) + up(sizeof(int) * (7 * ni + scan_tile_count(ni) + 1)) +
           up(sizeof(float) * QCS_REC_FLOATS * 2 * ni);
}

void cascade_carve(void *base, int n, CascadeBufs &b)
{
    auto up = [](size_t v) { return (v + 255) & ~(size_t)255; };
    const size_t ni = (size_t)(n > 0 ? n : 1);
    char *p = static_cast<char *>(base);
    b.cnt = reinterpret_cast<int *>(p);
    p += up(64);
    b.list1 = reinterpret_cast<int *>(p);
    p += up(sizeof(int) * ni);
    b.list2 = reinterpret_cast<int *>(p);
    p += up(sizeof(int) * ni);
    b.list3 = reinterpret_cast<int *>(p);
    p += up(sizeof(int) * ni);
    b.flag = reinterpret_cast<int *>(p);
    p += up(sizeof(int) * (7 * ni + scan_tile_count(ni) + 1));
    b.rows = reinterpret_cast<float *>(p);
}

int fit_cascade_launch(const qcs_params &p, Timer *tm, const uint16_t *d_rois, const float *d_wle, int n, float *d_loc,
                       int loc_cap, const CascadeBufs &b, hipStream_t s)
{
    const bool as3d = (p.LocType == QCS_LOCTYPE_AS3D);
    const int ntiles = scan_tile_count((size_t)n);
    int rc;
    int *list1 = b.list1, *list2 = b.list2, *list3 = b.list3;
    int *f1 = b.flag, *f2 = f1 + n, *rr = f2 + n, *fs2 = rr + n, *ff2 = fs2 + n, *fs3 = ff2 + n;
    int *offs = fs3 + n, *tsum = offs + n;
    int *ff3 = f1;   // f1 is free again once the 1E list is built
    float *sec2 = b.rows, *sec3 = sec2 + (size_t)QCS_REC_FLOATS * n;
    int *cnt = b.cnt;
    const dim3 g((n + 255) / 256);
    // stage timers are recorded only when a timer is given and enabled
    auto tb = [&](hipEvent_t *ev) { *ev = nullptr; return tm ? timer_begin(*tm, s, ev) : QCS_OK; };
    auto te = [&](int cat, hipEvent_t ev) { return tm ? timer_end(*tm, cat, s, ev, n) : QCS_OK; };

    fill_async(cnt, 16, 0, s);
    hipLaunchKernelGGL(cascade_set_kernel, dim3(1), dim3(1), 0, s, cnt + 8, loc_cap);
    fill_async(rr, n, 0, s);
    fill_async(fs2, n, 0, s);
    fill_async(ff2, n, 0, s);
    fill_async(fs3, n, 0, s);
    hipLaunchKernelGGL(cascade_classify_kernel, g, dim3(256), 0, s, d_wle, n, f1, f2);
    // 1E list and classified part of the 2E list
    if ((rc = scan_exclusive(f1, n, offs, tsum, s)) || (rc = compact_scatter(f1, offs, n, nullptr, nullptr, 0, list1, s)))
        return rc;
    hipLaunchKernelGGL(cascade_count_kernel, dim3(1), dim3(1), 0, s, tsum, ntiles, cnt, 0, 0);
    if ((rc = scan_exclusive(f2, n, offs, tsum, s)) || (rc = compact_scatter(f2, offs, n, nullptr, nullptr, 0, list2, s)))
        return rc;
    hipLaunchKernelGGL(cascade_count_kernel, dim3(1), dim3(1), 0, s, tsum, ntiles, cnt, 1, 0);
    hipLaunchKernelGGL(cascade_count_kernel, dim3(1), dim3(1), 0, s, tsum, ntiles, cnt, 2, 0);
    // one-emitter stage; rejected fits with SNR > 6 flag themselves in rr
    hipEvent_t ev;
    if ((rc = tb(&ev))) return rc;
    if ((rc = fit_launch_1e(p, d_rois, d_wle, list1, n, cnt + 0, d_loc, rr, s))) return rc;
    if ((rc = te(TM_FIT1E, ev))) return rc;
    // re-routed ROIs join the 2E list after the classified ones, in index order
    if ((rc = scan_exclusive(rr, n, offs, tsum, s)) || (rc = compact_scatter(rr, offs, n, nullptr, cnt + 1, 0, list2, s)))
        return rc;
    hipLaunchKernelGGL(cascade_count_kernel, dim3(1), dim3(1), 0, s, tsum, ntiles, cnt, 2, 1);
    // two-emitter stage
    if ((rc = tb(&ev))) return rc;
    if ((rc = fit_launch_me(p, as3d ? MODEL_AS3D2E : MODEL_2E, d_rois, d_wle, list2, n, cnt + 2, d_loc, sec2, fs2, ff2, s)))
        return rc;
    if ((rc = te(TM_FITME, ev))) return rc;
    hipLaunchKernelGGL(clear_tail_kernel, g, dim3(256), 0, s, fs2, n, cnt + 2);
    hipLaunchKernelGGL(clear_tail_kernel, g, dim3(256), 0, s, ff2, n, cnt + 2);
    if ((rc = scan_exclusive(fs2, n, offs, tsum, s))) return rc;
    hipLaunchKernelGGL(append_kernel, g, dim3(256), 0, s, fs2, offs, cnt + 2, sec2, cnt, -1, n, loc_cap, d_loc);
    hipLaunchKernelGGL(cascade_count_kernel, dim3(1), dim3(1), 0, s, tsum, ntiles, cnt, 4, 0);
    if ((rc = scan_exclusive(ff2, n, offs, tsum, s))) return rc;
    if (as3d) {
        hipLaunchKernelGGL(cascade_count_kernel, dim3(1), dim3(1), 0, s, tsum, ntiles, cnt, 6, 0);
    } else {
        if ((rc = compact_scatter(ff2, offs, n, list2, nullptr, 0, list3, s))) return rc;
        hipLaunchKernelGGL(cascade_count_kernel, dim3(1), dim3(1), 0, s, tsum, ntiles, cnt, 3, 0);
        fill_async(ff3, n, 0, s);
        if ((rc = tb(&ev))) return rc;
        if ((rc = fit_launch_me(p, MODEL_3E, d_rois, d_wle, list3, n, cnt + 3, d_loc, sec3, fs3, ff3, s))) return rc;
        if ((rc = te(TM_FITME, ev))) return rc;
        hipLaunchKernelGGL(clear_tail_kernel, g, dim3(256), 0, s, fs3, n, cnt + 3);
        hipLaunchKernelGGL(clear_tail_kernel, g, dim3(256), 0, s, ff3, n, cnt + 3);
        if ((rc = scan_exclusive(fs3, n, offs, tsum, s))) return rc;
        hipLaunchKernelGGL(append_kernel, g, dim3(256), 0, s, fs3, offs, cnt + 3, sec3, cnt, 4, n, loc_cap, d_loc);
        hipLaunchKernelGGL(cascade_count_kernel, dim3(1), dim3(1), 0, s, tsum, ntiles, cnt, 5, 0);
        if ((rc = scan_exclusive(ff3, n, offs, tsum, s))) return rc;
        hipLaunchKernelGGL(cascade_count_kernel, dim3(1), dim3(1), 0, s, tsum, ntiles, cnt, 6, 0);
    }
    hipLaunchKernelGGL(cascade_rows_kernel, dim3(1), dim3(1), 0, s, n, cnt);
    QCS_HIP(hipGetLastError());
    if ((rc = tb(&ev))) return rc;
    if ((rc = crlb_launch(p, d_loc, loc_cap, cnt + 7, s))) return rc;
    return te(TM_CRLB, ev);
}

int fit_cascade(qcs_ctx *c, const uint16_t *d_rois, const float *d_wle, int n, float *d_loc, int loc_cap, int *n_rows,
                hipStream_t s)
{
    const qcs_params &p = c->p;
    const bool as3d = (p.LocType == QCS_LOCTYPE_AS3D);
    int rc;
    if ((rc = c->fit_flag.ensure(cascade_scratch_bytes(n))) || (rc = c->pin_cnt.ensure(64))) return rc;
    CascadeBufs b;
    cascade_carve(c->fit_flag.ptr, n, b);
    if ((rc = fit_cascade_launch(p, &c->timer, d_rois, d_wle, n, d_loc, loc_cap, b, s))) return rc;
    int *cnt = b.cnt;
    int *h = c->pin_cnt.as<int>();
    QCS_HIP(hipMemcpyAsync(h, cnt, 16 * sizeof(int), hipMemcpyDeviceToHost, s));
    QCS_HIP(hipStreamSynchronize(s));
    qcs_fit_stats &st = c->stats;
    st.n_in = n;
    st.n_single = h[0];
    st.n_2e = h[2];
    st.n_3e = h[3];
    st.n_rejected = h[6];
    st.n_rows = h[7];
    // FitRatio_* as bfgs_top.cu:64-173
    const float nd = (float)(n < 1 ? 1 : n);
    st.FitRatio_1E = (float)h[0] / nd;
    st.FitRatio_2E = (float)h[2] / nd;
    st.FitRatio_3E = (float)h[3] / nd;
    st.FitRatio_Final_1E = (float)(h[0] - (h[2] - h[1])) / nd;
    if (!as3d) {
        st.FitRatio_Final_2E = (float)(h[2] - h[3]) / nd;
        st.FitRatio_Final_3E = (float)(h[3] - h[6]) / nd;
    } else {
        st.FitRatio_Final_2E = (float)(h[2] - h[6]) / nd;
        st.FitRatio_Final_3E = 0.0f;
    }
    st.FitRatio_Final_4E = (float)h[6] / nd;
    if (n_rows) *n_rows = h[7];
    return QCS_OK;
}

}  // namespace qcs
