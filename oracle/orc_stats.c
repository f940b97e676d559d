/*
 * orc_stats.c -- statistics of a batch of localization rows, restated for the CPU oracle
 * (TEST INFRASTRUCTURE).
 *
 * Reference (DLL = MLEbfgs/bfgs_MLE_dll x64/bfgsMLE_dll/bfgsMLE_dll):
 *   DLL/StatisticsInfo/statisticsInfo.h:40-62          histogram sizes and bin widths, radii
 *   DLL/StatisticsInfo/statisticsInfo.cu:19-91         CalcLocResultStatInf (filter, histograms, sums)
 *   DLL/StatisticsInfo/statisticsInfo_top.cu:61-128    GetStatisticalInf
 *   DLL/StatisticsInfo/statisticsInfo_top.cu:130-205   CalcLocalizationDensity2D
 *   DLL/StatisticsInfo/statisticsInfo_top.cu:208-225   UpdateStatDat (means)
 *   DLL/StatisticsInfo/statisticsInfo_densityCalc.cu:17-107  frame filter, neighbour counts
 *   DLL/StatisticsInfo/statisticsInfo_rc.cu:17-74      GetActivationDensity, GetHistogramMeanData
 *   DLL/bfgsMLE_Loc_top/bfgs_top.cu:345-381            GetFirstFrame, GetLastFrame
 * The reference sums with float atomics (order-dependent); the oracle sums in double in row
 * order.  Neighbour counts are the reference's brute force over the molecules of each frame.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "orc.h"

#define HIST_LEN   400
#define HIST_MAX   (HIST_LEN * 4)

/* bin of value / width clamped to [0, HIST_MAX - 1]; the reference's float-to-int conversion
 * saturates (a CUDA cvt), so the clamp is applied to the float before converting */
static int clamp_bin(float q) { return q >= (float)HIST_MAX ? HIST_MAX - 1 : (q > 0.0f ? (int)q : 0); }

/* statisticsInfo_rc.cu:17-41 */
float orc_activation_density(float r0, float r1, float radius_um)
{
    const float PI = 3.14159f;
    if (r0 <= 0.001f) r0 = 0.001f;
    if (r1 <= 0.001f) r1 = 0.001f;
    if (r0 > 1) r0 = 1;
    if (r1 > 1) r1 = 1;
    const float l0 = logf(r0);
    const float l1 = logf(r1 / PI / radius_um / radius_um);
    const float d0 = -l0 / PI / radius_um / radius_um;
    const float d1 = -l1 / PI / radius_um / radius_um;
    return r0 * d0 + (1.0f - r0) * d1;
}

/* statisticsInfo_rc.cu:44-74 (GetHistogramMeanData with GetHistogramMaxData) */
float orc_hist_mean(const int *h, int len, float pct)
{
    int mx = h[0];
    for (int i = 0; i < len; i++) mx = h[i] > mx ? h[i] : mx;
    pct = pct * (float)mx;
    float ws = 0.0f, os = 0.0f;
    for (int i = 0; i < len; i++) {
        const float v = (float)h[i];
        if (v >= pct) {
            ws += v * ((float)i + 0.5f);
            os += v;
        }
    }
    if (os == 0) os = 1;
    return ws / os;
}

/* GetFirstFrame / GetLastFrame (bfgs_top.cu:345-381): frame of the first row with a non-zero
 * frame; frame of the last such row among rows n-1 .. 1 (row 0 is not examined) */
static int first_frame(const float *loc, int n)
{
    int f = 0;
    for (int i = 0; i < n; i++) {
        f = (int)loc[(size_t)i * QCS_REC_FLOATS + QCS_POS_FRME];
        if (f != 0) break;
    }
    return f;
}

static int last_frame(const float *loc, int n)
{
    int f = 0;
    for (int i = n - 1; i > 0; i--) {
        f = (int)loc[(size_t)i * QCS_REC_FLOATS + QCS_POS_FRME];
        if (f != 0) break;
    }
    return f;
}

int orc_loc_stats(const qcs_params *p, const float *loc, int n, qcs_loc_stats *o)
{
    memset(o, 0, sizeof(*o));
    if (n <= 0) return 0;
    /* CalcLocResultStatInf (statisticsInfo.cu:19-91) */
    for (int i = 0; i < n; i++) {
        const float *r = loc + (size_t)i * QCS_REC_FLOATS;
        const float tp = r[QCS_POS_TPHO], bg = r[QCS_POS_BAKG];
        const float sg = (r[QCS_POS_SIGX] + r[QCS_POS_SIGY]) / 2;
        const float snr = r[QCS_POS_PSNR];
        const float cx = r[QCS_POS_CRBX], cy = r[QCS_POS_CRBY];
        const float lp = sqrtf(cx * cx + cy * cy);
        if (!((tp > 30.0f) && (bg > 2.0f) && (snr >= 3.0f) && (sg >= 0.01f) && (lp > 0.01f) && (lp < 55.0))) continue;
        o->hist_total_photon[clamp_bin(tp / (10000.0f / HIST_LEN))]++;
        o->hist_loc_prec[clamp_bin(lp / (60.0f / HIST_LEN))]++;
        o->hist_peak_snr[clamp_bin(snr / (50.0f / HIST_LEN))]++;
        o->hist_psf_width[clamp_bin(sg / (5.0f / HIST_LEN))]++;
        o->valid++;
        o->sum_loc_prec += lp;
        o->sum_peak_snr += snr;
        o->sum_psf_width += sg;
        o->sum_background += bg;
    }
    /* UpdateStatDat (statisticsInfo_top.cu:208-225) */
    const double vn = o->valid > 0 ? (double)o->valid : 1.0;
    o->mean_total_photon = (10000.0f / HIST_LEN) * orc_hist_mean(o->hist_total_photon, HIST_MAX, 0.1f);
    o->mean_loc_prec = (float)(o->sum_loc_prec / vn);
    o->mean_peak_snr = (float)(o->sum_peak_snr / vn);
    o->mean_background = (float)(o->sum_background / vn);
    o->mean_psf_width = (float)(o->sum_psf_width / vn);
    o->mean_psf_width_ctl = o->mean_psf_width * (-o->mean_peak_snr);

    /* CalcLocalizationDensity2D (statisticsInfo_top.cu:130-205) */
    const float radius = (p->LocType == QCS_LOCTYPE_GS2D) ? 10.0f : 20.0f;
    const float radius_um = radius * p->PixelSize / 1000;
    int fs = first_frame(loc, n), fe = last_frame(loc, n);
    if (fe - fs + 1 > 3) {
        fs++;
        fe--;
    }
    o->density_first_frame = fs;
    o->density_last_frame = fe;
    int *sel = (int *)malloc(sizeof(int) * (size_t)n);
    const float r2 = radius * radius;
    for (int f = fs; f <= fe; f++) {
        /* FillerMoleculesInCurFrame (statisticsInfo_densityCalc.cu:17-57) */
        int m = 0;
        for (int i = 0; i < n; i++) {
            const float *r = loc + (size_t)i * QCS_REC_FLOATS;
            const int x = (int)r[QCS_POS_XPOS], y = (int)r[QCS_POS_YPOS], cf = (int)r[QCS_POS_FRME];
            if (x >= 5 && x < p->ImageWidth - 5 && y >= 5 && y < p->ImageHigh - 5 && cf == f) sel[m++] = i;
        }
        /* CalcOverlapFluoNum (statisticsInfo_densityCalc.cu:60-104) */
        for (int a = 0; a < m; a++) {
            const float *ra = loc + (size_t)sel[a] * QCS_REC_FLOATS;
            const float ax = ra[QCS_POS_XPOS], ay = ra[QCS_POS_YPOS];
            int cnt = 0;
            for (int b = 0; b < m; b++) {
                const float *rb = loc + (size_t)sel[b] * QCS_REC_FLOATS;
                const float dx = rb[QCS_POS_XPOS] - ax, dy = rb[QCS_POS_YPOS] - ay;
                const float d2 = dx * dx + dy * dy + (float)(a == b) * 1000000.0f;
                cnt += d2 <= r2;
            }
            if (cnt == 0) o->density_n0++;
            if (cnt == 1) o->density_n1++;
        }
        o->density_total += m;
    }
    free(sel);
    if (o->density_total > 50) {
        const float r0 = (float)o->density_n0 / (float)o->density_total, r1 = (float)o->density_n1 / (float)o->density_total;
        float d = orc_activation_density(r0, r1, radius_um);
        d = p->MultiEmitterFitEn ? d * d * 0.5098f + d * 0.6364f : d * d * 0.8082f + d * 0.519f;
        o->loc_density_2d = d;
    }
    return 0;
}
