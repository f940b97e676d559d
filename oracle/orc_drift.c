/*
 * orc_drift.c -- drift correction by cross-correlation of super-resolution images, restated for
 * the CPU oracle (TEST INFRASTRUCTURE).
 *
 * Reference (DLL = MLEbfgs/bfgs_MLE_dll x64/bfgsMLE_dll/bfgsMLE_dll/DriftCorrection_SRImageXCorr):
 *   DLL/DriftCorrection_SRImageXCorr.h:31-48             CorrSize 47, FittingRadius 12, zoom 8,
 *                                                        7x7 render window, MaxCorrGroupNum 512
 *   DLL/DriftCorrection_SRImageXCorr_top.cu:18-74        CorrectSampleShift
 *   DLL/DriftCorrection_SRImageXCorr_top.cu:172-298      GetSliceShift (peak, centre of mass)
 *   DLL/DriftCorrection_SRImageXCorr_rc.cu:20-73         ShiftInterpolation
 *   DLL/DriftCorrection_SRImageXCorr_rc.cu:170-208       CalcGroupNum
 *   DLL/DriftCorrection_SRImageXCorr_rc.cu:211-287       GetCorrGroupFluoPos, GetAFrameEndFluoPos
 *   DLL/DriftCorrection_SRImageXCorr_rc.cu:322-356       CrossCorrelation (column sums, double total)
 *   DLL/DriftCorrection_SRImageXCorr_rc.cu:360-395       GetTotalFrame
 *   DLL/DriftCorrection_SRImageXCorr.cu:21-125           CreateRenderPSF, PointFillRegionWithPrec1
 *   DLL/DriftCorrection_SRImageXCorr.cu:127-157          gpuApplyShiftCorr
 *   DLL/DriftCorrection_SRImageXCorr.cu:159-212          gpuCrossCorr_Mul, gpuCrossCorr_SumOfMul
 * The render uses the oracle's deterministic exp (the reference's __expf is not reproducible) and
 * sums in row order; the correlation is the reference's: per shift, per column a float sum over
 * the rows in order, the columns summed in double.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "orc.h"

#define CORR     47
#define HALF     23
#define FITR     12
#define DZOOM    8
#define RW       7
#define RW_HALF  3
#define MAXGRP   512

int orc_drift_total_frame(const float *loc, long long n)
{
    const long long sel = n < 400 ? n : 400;
    for (long long i = n - 1; i >= n - sel; i--) {
        const int f = (int)loc[i * QCS_REC_FLOATS + QCS_POS_FRME];
        if (f > 0) return f;
    }
    return 0;
}

/* CalcGroupNum: groups of corr frames overlapping by half; returns the group count (<= 0 when
 * total < corr, which the reference does not handle) */
int orc_drift_groups(int total, int corr, int *gs, int *ge)
{
    const int half = corr / 2;
    int g = total / corr;
    g = g * 2 - 1;
    if (g <= 0 || g > MAXGRP - 1) return -1;
    const int resid = total % corr;
    for (int k = 0; k < g; k++) {
        gs[k] = k * half + 1;
        ge[k] = k * half + corr;
    }
    if (resid < half) {
        ge[g - 1] = total;
    } else {
        g++;
        gs[g - 1] = total - corr + 1;
        ge[g - 1] = total;
    }
    return g;
}

/* GetAFrameEndFluoPos: the row before the first row at or after `offset` whose frame is > 0 and
 * > find; n - 1 when there is none (or when that row is row 1 of the rows, the reference's
 * EndPos == 0 test) */
static long long frame_end_pos(const float *loc, long long n, long long offset, int find)
{
    long long end = 0;
    for (long long i = offset; i < n; i++) {
        const int f = (int)loc[i * QCS_REC_FLOATS + QCS_POS_FRME];
        if (f > 0 && f > find) {
            end = i - 1;
            break;
        }
    }
    if (end == 0) end = n - 1;
    return end;
}

void orc_drift_group_rows(const float *loc, long long n, int ng, const int *gs, const int *ge, long long *first, long long *last)
{
    long long b = 0, e = 0;
    first[0] = 0;
    for (int k = 0; k < ng; k++) {
        e = frame_end_pos(loc, n, e, ge[k]);
        last[k] = e;
        if (k >= 1) {
            b = frame_end_pos(loc, n, b, gs[k] - 1);
            first[k] = b + 1;
        }
    }
}

/* PointFillRegionWithPrec1 at zoom 8: unit weight, sigma from CRLBx (35 nm when outside
 * (0.01, 35)), rows with PSNR >= SNR_th, PPho > 1 and the window inside the image */
void orc_drift_render(const qcs_params *p, const float *loc, long long n, int srw, int srh, float *img)
{
    const float zoom = (float)DZOOM;
    for (long long i = 0; i < n; i++) {
        const float *r = loc + i * QCS_REC_FLOATS;
        const float pp = r[QCS_POS_PPHO];
        const float rx = r[QCS_POS_XPOS] * zoom, ry = r[QCS_POS_YPOS] * zoom;
        const float snr = r[QCS_POS_PSNR];
        float prec = r[QCS_POS_CRBX];
        if (!((snr >= p->SNR_th) && (pp > 1.0f) && (rx > RW) && (rx < srw - 1 - RW) && (ry > RW) && (ry < srh - 1 - RW))) continue;
        if (!((prec > 0.01f) && (prec < 35.0f))) prec = 35.0f;
        const float sig = prec / (p->PixelSize / zoom);
        const float rs = -0.5f / (sig * sig);
        float fx = rx - (float)(int)rx, fy = ry - (float)(int)ry;
        fx = fx + (float)RW_HALF;
        fy = fy + (float)RW_HALF;
        for (int yy = 0; yy < RW; yy++) {
            const long long off = (long long)srw * ((int)ry - RW_HALF + yy) + (int)rx - RW_HALF;
            for (int xx = 0; xx < RW; xx++) {
                const float dx = (float)xx - fx, dy = (float)yy - fy;
                img[off + xx] += 1.0f * orc_qexpf(fmaf(dx, dx, dy * dy) * rs);
            }
        }
    }
}

/* CrossCorrelation for one shift (including the bias): sum over img1 pixels p of
 * img1[p] * img2[p + shift] where p + shift is inside the image; per column a float sum over the
 * rows in order, the columns summed in double */
double orc_drift_cc(const float *img1, const float *img2, int w, int h, int sx, int sy)
{
    double tot = 0.0;
    for (int x = 0; x < w; x++) {
        float s = 0.0f;
        const int x1 = x + sx;
        if (x1 >= 0 && x1 < w)
            for (int y = 0; y < h; y++) {
                const int y1 = y + sy;
                if (y1 >= 0 && y1 < h) s += img1[(size_t)y * w + x] * img2[(size_t)y1 * w + x1];
            }
        tot += s;
    }
    return tot;
}

/* GetSliceShift after the correlation values: the first maximum, normalisation to 1000, centre of
 * mass of the row and column sums over the fitting window; the shift in SR pixels */
void orc_drift_peak(const double *cc, int bias_x, int bias_y, float *shift_x, float *shift_y)
{
    double c[CORR * CORR];
    double mx = 0.0;
    int mxx = 0, mxy = 0;
    for (int y = 0; y < CORR; y++)
        for (int x = 0; x < CORR; x++) {
            c[y * CORR + x] = cc[y * CORR + x];
            if (mx < c[y * CORR + x]) {
                mx = c[y * CORR + x];
                mxx = x;
                mxy = y;
            }
        }
    for (int k = 0; k < CORR * CORR; k++) c[k] = c[k] / mx * 1000;
    const int xs = mxx - FITR > 0 ? mxx - FITR : 0, ys = mxy - FITR > 0 ? mxy - FITR : 0;
    const int xe = mxx + FITR < CORR - 1 ? mxx + FITR : CORR - 1, ye = mxy + FITR < CORR - 1 ? mxy + FITR : CORR - 1;
    const int xh = (xe - xs + 1) / 2, yh = (ye - ys + 1) / 2;
    float sx[2 * FITR + 1], sy[2 * FITR + 1];
    for (int k = 0; k < 2 * FITR + 1; k++) sx[k] = sy[k] = 0.0f;
    for (int y = ys; y <= ye; y++)
        for (int x = xs; x <= xe; x++) {
            sx[FITR - xh + (x - xs)] = (float)((double)sx[FITR - xh + (x - xs)] + c[y * CORR + x]);
            sy[FITR - yh + (y - ys)] = (float)((double)sy[FITR - yh + (y - ys)] + c[y * CORR + x]);
        }
    float ws = 0, cs = 0;
    for (int k = 0; k < 2 * FITR + 1; k++) {
        if (sx[k] == 0) continue;
        ws += sx[k] * (float)(k + 1);
        cs += sx[k];
    }
    const float cx = ws / cs - (float)(FITR + 1);
    ws = 0;
    cs = 0;
    for (int k = 0; k < 2 * FITR + 1; k++) {
        if (sy[k] == 0) continue;
        ws += sy[k] * (float)(k + 1);
        cs += sy[k];
    }
    const float cy = ws / cs - (float)(FITR + 1);
    *shift_x = (float)(bias_x + mxx - HALF) + cx;
    *shift_y = (float)(bias_y + mxy - HALF) + cy;
}

/* ShiftInterpolation: per-frame shifts (index frame - 1), linear between group centres, from 0 at
 * frame index 0, the last slope continued to the last group's end frame */
void orc_drift_interp(int total, int ng, const int *gs, const int *ge, const float *gx, const float *gy, float *fx, float *fy)
{
    memset(fx, 0, sizeof(float) * (size_t)total);
    memset(fy, 0, sizeof(float) * (size_t)total);
    float lc = 0, lx = 0, ly = 0, cc = 0, cx = 0, cy = 0, kx = 0, ky = 0;
    for (int k = 0; k <= ng; k++) {
        if (k == ng) {
            cc = (float)ge[ng - 1];
        } else {
            cc = (float)((gs[k] + ge[k]) / 2);
            cx = gx[k];
            cy = gy[k];
            kx = (cx - lx) / (cc - lc);
            ky = (cy - ly) / (cc - lc);
        }
        for (int f = (int)lc; (float)f < cc; f++) {
            fx[f] = lx + ((float)f - lc) * kx;
            fy[f] = ly + ((float)f - lc) * ky;
        }
        lc = cc;
        lx = cx;
        ly = cy;
    }
}

/* CorrectSampleShift on rows in memory: returns the group count (< 0 on error); rows are
 * corrected in place; fx/fy (total entries) receive the per-frame shifts; gx/gy the group shifts */
int orc_drift_correct(const qcs_params *p, float *loc, long long n, int corr, float *fx, float *fy, int fcap, int *total_out,
                      float *gx_out, float *gy_out)
{
    const int total = orc_drift_total_frame(loc, n);
    *total_out = total;
    if (total > fcap) return -2;
    int gs[MAXGRP], ge[MAXGRP];
    const int ng = orc_drift_groups(total, corr, gs, ge);
    if (ng <= 0) return -1;
    long long first[MAXGRP], last[MAXGRP];
    orc_drift_group_rows(loc, n, ng, gs, ge, first, last);
    const int srw = p->ImageWidth * DZOOM, srh = p->ImageHigh * DZOOM;
    float *img1 = (float *)calloc((size_t)srw * srh, sizeof(float));
    float *img2 = (float *)calloc((size_t)srw * srh, sizeof(float));
    double *cc = (double *)malloc(sizeof(double) * CORR * CORR);
    float gx[MAXGRP], gy[MAXGRP];
    float sxv = 0, syv = 0;
    int bx = 0, by = 0;
    for (int k = 0; k < ng; k++) {
        float *im = k == 0 ? img1 : img2;
        memset(im, 0, sizeof(float) * (size_t)srw * srh);
        orc_drift_render(p, loc + first[k] * QCS_REC_FLOATS, last[k] - first[k] + 1, srw, srh, im);
        if (k > 0) {
            for (int y = 0; y < CORR; y++)
                for (int x = 0; x < CORR; x++) cc[y * CORR + x] = orc_drift_cc(img1, img2, srw, srh, bx + x - HALF, by + y - HALF);
            orc_drift_peak(cc, bx, by, &sxv, &syv);
            bx = (int)sxv;
            by = (int)syv;
        }
        gx[k] = sxv / DZOOM;
        gy[k] = syv / DZOOM;
    }
    orc_drift_interp(total, ng, gs, ge, gx, gy, fx, fy);
    for (long long i = 0; i < n; i++) {
        float *r = loc + i * QCS_REC_FLOATS;
        const int f = (int)r[QCS_POS_FRME];
        if (f < 1 || f > total) continue;
        r[QCS_POS_XPOS] = r[QCS_POS_XPOS] - fx[f - 1];
        r[QCS_POS_YPOS] = r[QCS_POS_YPOS] - fy[f - 1];
    }
    if (gx_out)
        for (int k = 0; k < ng; k++) gx_out[k] = gx[k];
    if (gy_out)
        for (int k = 0; k < ng; k++) gy_out[k] = gy[k];
    free(img1);
    free(img2);
    free(cc);
    return ng;
}
