/*
 * orc_wle.c -- weighted-likelihood (WLE) pre-estimation, restated for the CPU oracle
 * (TEST INFRASTRUCTURE).
 *
 * Reference (DLL = MLEbfgs/bfgs_MLE_dll x64/bfgsMLE_dll/bfgsMLE_dll):
 *   DLL/WLEParaEstimation/WLEParaEstimation_core.h:44-156   PSF width per ROI
 *   DLL/WLEParaEstimation/SplineInterpolation.h:40-202      natural cubic spline (Thomas)
 *   DLL/WLEParaEstimation/WLEParaEstimation_core.cu:91-136  FindMaxPos, SigmaWidthCalc_TwoSides
 *   DLL/WLEParaEstimation/WLEParaEstimation_core.cu:140-183 nearest-neighbour distance
 *   DLL/WLEParaEstimation/WLEParaEstimation_core.cu:197-264 molecule type
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "orc.h"

#define MAXLEN 19
#define MAXINT ((MAXLEN - 1) * 10 + 1)

/* Natural cubic spline through (x_i, y_i), evaluated at x_o (SplineInterpolation,
 * SplineInterpolation.h:40-74; system :103-146; Thomas solve :151-202; evaluation :79-101 with
 * FindPositionID_0 = truncation of x, WLEParaEstimation_core.cu:21-28). */
void orc_spline_interp(const float *x_i, const float *y_i, int n, const float *x_o, float *y_o, int n_out)
{
    float A[MAXLEN][MAXLEN], b[MAXLEN], cn[MAXLEN], bn[MAXLEN], dn[MAXLEN];
    float beta[MAXLEN], yv[MAXLEN];
    int i, r, c;
    for (r = 0; r < n; r++)
        for (c = 0; c < n; c++) A[r][c] = 0.0f;
    A[0][0] = 1.0f;
    A[n - 1][n - 1] = 1.0f;
    for (i = 1; i < n - 1; i++) {
        float h0 = x_i[i] - x_i[i - 1];
        float h1 = x_i[i + 1] - x_i[i];
        A[i][i - 1] = h0;
        A[i][i] = 2.0f * (h0 + h1);
        A[i][i + 1] = h1;
    }
    b[0] = 0.0f;
    b[n - 1] = 0.0f;
    for (i = 1; i < n - 1; i++) {
        float h0 = x_i[i] - x_i[i - 1];
        float h1 = x_i[i + 1] - x_i[i];
        b[i] = fmaf(3.0f / h1, y_i[i + 1] - y_i[i], -((3.0f / h0) * (y_i[i] - y_i[i - 1])));
    }
    /* forward elimination factors */
    for (i = 0; i < n - 1; i++) {
        float ci = A[i][i + 1], bi = A[i][i];
        if (i == 0) beta[i] = ci / bi;
        else beta[i] = ci / fmaf(-A[i][i - 1], beta[i - 1], bi);
    }
    for (i = 0; i < n; i++) {
        float bi = A[i][i];
        if (i == 0) yv[i] = b[i] / bi;
        else yv[i] = fmaf(-A[i][i - 1], yv[i - 1], b[i]) / fmaf(-A[i][i - 1], beta[i - 1], bi);
    }
    for (i = n - 1; i >= 0; i--) {
        if (i == n - 1) cn[i] = yv[i];
        else cn[i] = fmaf(-beta[i], cn[i + 1], yv[i]);
    }
    for (i = 0; i < n - 1; i++) {
        float hi = x_i[i + 1] - x_i[i];
        bn[i] = fmaf(1.0f / hi, y_i[i + 1] - y_i[i], -((hi / 3.0f) * fmaf(2.0f, cn[i], cn[i + 1])));
        dn[i] = (cn[i + 1] - cn[i]) / (3.0f * hi);
    }
    for (i = 0; i < n_out; i++) {
        float xn = x_o[i];
        int ps = (int)(xn - 0.0f);
        if (ps < 0) ps = 0;
        if (ps >= n - 1) ps = n - 2;
        float t = xn - x_i[ps];
        float v = fmaf(bn[ps], t, y_i[ps]);
        v = fmaf(cn[ps] * t, t, v);
        v = fmaf((dn[ps] * t) * t, t, v);
        y_o[i] = v;
    }
}

/* FindMaxPos (WLEParaEstimation_core.cu:120-136) */
static int find_max_pos(const float *d, int len, int lsel, int rsel)
{
    if (lsel < 0) lsel = 0;
    if (rsel > len - 1) rsel = len - 1;
    int mx = len / 2;
    for (int c = lsel; c <= rsel; c++)
        if (d[c] > d[mx]) mx = c;
    return mx;
}

/* SigmaWidthCalc_TwoSides (WLEParaEstimation_core.cu:91-118) */
static void sigma_two_sides(float *sl, float *sr, const float *d, int len, int mp)
{
    float mv = d[mp];
    int ll = mp, rl = len - mp - 1;
    int cl = ll < rl ? ll : rl;
    float a = 0.0f, b = 0.0f;
    for (int i = mp - cl; i <= mp; i++) a = a + d[i];
    for (int i = mp; i <= mp + cl; i++) b = b + d[i];
    *sl = ((a * 0.1f) / mv) / 1.2533f;
    *sr = ((b * 0.1f) / mv) / 1.2533f;
}

/* gpu_CalculatePSFWidth<ROISize> (WLEParaEstimation_core.h:44-156), MarginLen = 0 */
void orc_wle_psf_width(const uint16_t *rois, int n, int R, float *wle)
{
    const int W = QCS_ROI_WORDS(R);
    const int nint = (R - 1) * 10 + 1;
    float x0[MAXLEN], xn[MAXINT], mx[MAXLEN], my[MAXLEN], ymx[MAXINT], ymy[MAXINT];
    int i, r, c;
    for (c = 0; c < R; c++) x0[c] = (float)c;
    for (c = 0; c < nint; c++) xn[c] = (float)c * 0.1f;
    for (i = 0; i < n; i++) {
        const uint16_t *roi = rois + (long)i * W;
        for (c = 0; c < R; c++) { mx[c] = 0.0f; my[c] = 0.0f; }
        for (r = 0; r < R; r++)
            for (c = 0; c < R; c++) {
                mx[c] = mx[c] + (float)roi[r * R + c];
                my[r] = my[r] + (float)roi[r * R + c];
            }
        float mnx = mx[0], mny = my[0];
        for (c = 0; c < R; c++) {
            mnx = fminf(mnx, mx[c]);
            mny = fminf(mny, my[c]);
        }
        for (c = 0; c < R; c++) {
            mx[c] = (mx[c] - mnx) / (float)R;
            my[c] = (my[c] - mny) / (float)R;
        }
        orc_spline_interp(x0, mx, R, xn, ymx, nint);
        orc_spline_interp(x0, my, R, xn, ymy, nint);
        const int region = 8;            /* (int)(InterpolationRatio * PeakFindRangeRatio) */
        const int ctr = nint / 2;
        int px = find_max_pos(ymx, nint, ctr - region, ctr + region);
        int py = find_max_pos(ymy, nint, ctr - region, ctr + region);
        float sl, sr, su, sd;
        sigma_two_sides(&sl, &sr, ymx, nint, px);
        sigma_two_sides(&su, &sd, ymy, nint, py);
        float *o = wle + (long)i * QCS_WLE_FLOATS;
        o[QCS_WLE_SIGMAL] = sl;
        o[QCS_WLE_SIGMAR] = sr;
        o[QCS_WLE_SIGMAU] = su;
        o[QCS_WLE_SIGMAD] = sd;
    }
}

/* gpu_CalculateNearestNeighborDistance (WLEParaEstimation_core.cu:140-183): brute force over
 * all ROIs; other frames and self carry a 1e8 bias; result is sqrt of the minimum. */
void orc_wle_nearest(const uint16_t *rois, int n, int R, float *wle)
{
    const int W = QCS_ROI_WORDS(R), P = R * R;
    for (int i = 0; i < n; i++) {
        const uint16_t *h = rois + (long)i * W + P;
        int cx = h[0], cy = h[1], cf = (int)h[2] + (int)h[3] * 65536;
        float best = 100000000.0f;
        for (int j = 0; j < n; j++) {
            const uint16_t *g = rois + (long)j * W + P;
            int nx = g[0], ny = g[1], nf = (int)g[2] + (int)g[3] * 65536;
            int bias = (int)((float)(nf != cf) * 100000000.0f + (float)(j == i) * 100000000.0f);
            float d = (float)((cx - nx) * (cx - nx) + (cy - ny) * (cy - ny) + bias);
            best = fminf(best, d);
        }
        wle[(long)i * QCS_WLE_FLOATS + QCS_WLE_NEARDIST] = sqrtf(best);
    }
}

/* gpu_MoleculeTypeClasify (WLEParaEstimation_core.cu:197-264) */
void orc_wle_classify(const uint16_t *rois, int n, int R, int loc_type, float *wle)
{
    const int W = QCS_ROI_WORDS(R), P = R * R;
    const float th = (loc_type == QCS_LOCTYPE_GS2D) ? ((float)R / 2.1f) / 2.35f : ((float)R / 1.9f) / 2.35f;
    for (int i = 0; i < n; i++) {
        float *o = wle + (long)i * QCS_WLE_FLOATS;
        int det = rois[(long)i * W + P + 4];
        float nn = o[0], sl = o[1], sr = o[2], su = o[3], sd = o[4];
        float d1 = fabsf(sl - sr) / fminf(sl, sr);
        float d2 = fabsf(su - sd) / fminf(su, sd);
        int contam = (d1 >= 0.40f) || (d2 >= 0.40f);
        float ms = (((sl + sr) + su) + sd) / 4.0f;
        int mt;
        if (det == QCS_ROITYPE_SINGLE && contam == 0 && nn >= (float)R && ms <= th)
            mt = QCS_MOLTYPE_MLEFIT;
        else if (det == QCS_ROITYPE_MULTI || ms > th || nn < ((float)R / 2.0f + 1.0f))
            mt = QCS_MOLTYPE_MULTIFIT;
        else
            mt = QCS_MOLTYPE_WLEFIT;
        o[QCS_WLE_FLUOTYPE] = (float)mt;
    }
}

/* WLEParameterEstimation_TypeDef::WLEParameterEstimate (WLEParaEstimation.cpp:20-40) */
void orc_wle(const uint16_t *rois, int n, int R, int loc_type, float *wle)
{
    orc_wle_psf_width(rois, n, R, wle);
    orc_wle_nearest(rois, n, R, wle);
    orc_wle_classify(rois, n, R, loc_type, wle);
}
