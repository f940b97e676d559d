/*
 * orc_post.c -- frame-sorted compaction and splat rendering, restated for the CPU oracle
 * (TEST INFRASTRUCTURE).
 *
 * Reference (DLL = MLEbfgs/bfgs_MLE_dll x64/bfgsMLE_dll/bfgsMLE_dll):
 *   DLL/bfgsMLE_Loc_top/ZeroLocalizationsRemove.cu:42-61,161-186  sorted compaction
 *   DLL/bfgsMLE_Loc_top/ZeroLocalizationsRemove.cu:64-102,142-159  unsorted compaction
 *   DLL/ImgRender/ImageRender_2d.cu:25-47   CreateRenderPSF
 *   DLL/ImgRender/ImageRender_2d.cu:55-155  FluoRenderWithPrec_2D
 *   DLL/ImgRender/ImageRender_3d.cu:60-147  FluoRenderWithPrec_3D
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "orc.h"

float orc_exp32(float x);

/* (int)v as the GPU conversion does it: NaN -> 0, saturating */
static int f2i_sat(float v)
{
    if (v != v) return 0;
    if (v >= 2147483647.0f) return 2147483647;
    if (v <= -2147483648.0f) return (-2147483647 - 1);
    return (int)v;
}

/* For frames first..end, rows with (int)PPho > 1 and (int)Frame == frame, in frame order; the
 * reference's within-frame order comes from an atomic counter, the oracle keeps row order. */
int orc_compact_sorted(const float *loc, int n, int first, int end, float *out)
{
    int m = 0;
    if (end < first) return 0;
    int nf = end - first + 1;
    int *cnt = (int *)calloc((size_t)nf + 1, sizeof(int));
    for (int i = 0; i < n; i++) {
        const float *r = loc + (long)i * QCS_REC_FLOATS;
        int pp = f2i_sat(r[QCS_POS_PPHO]), fr = f2i_sat(r[QCS_POS_FRME]);
        if ((float)pp > 1.0f && fr >= first && fr <= end) cnt[fr - first + 1]++;
    }
    for (int f = 0; f < nf; f++) cnt[f + 1] += cnt[f];
    for (int i = 0; i < n; i++) {
        const float *r = loc + (long)i * QCS_REC_FLOATS;
        int pp = f2i_sat(r[QCS_POS_PPHO]), fr = f2i_sat(r[QCS_POS_FRME]);
        if ((float)pp > 1.0f && fr >= first && fr <= end) {
            memcpy(out + (long)cnt[fr - first] * QCS_REC_FLOATS, r, sizeof(float) * QCS_REC_FLOATS);
            cnt[fr - first]++;
            m++;
        }
    }
    free(cnt);
    return m;
}

/* FindCopyID + gpuRemoveZeorLocalizations: PPho > 1.0f, original order */
int orc_compact(const float *loc, int n, float *out)
{
    int m = 0;
    for (int i = 0; i < n; i++) {
        const float *r = loc + (long)i * QCS_REC_FLOATS;
        if (r[QCS_POS_PPHO] > 1.0f) {
            memcpy(out + (long)m * QCS_REC_FLOATS, r, sizeof(float) * QCS_REC_FLOATS);
            m++;
        }
    }
    return m;
}

/* CreateRenderPSF: 7x7 Gaussian centred at frac(x)+3, frac(y)+3 */
static void render_psf(float *psf, float rx, float ry, float rs)
{
    rs = -0.5f / (rs * rs);
    rx = rx - (float)(int)rx;
    ry = ry - (float)(int)ry;
    rx = rx + 3.0f;
    ry = ry + 3.0f;
    for (int y = 0; y < 7; y++)
        for (int x = 0; x < 7; x++) {
            float dx = (float)x - rx, dy = (float)y - ry;
            psf[y * 7 + x] = orc_exp32(fmaf(dx, dx, dy * dy) * rs);
        }
}

void orc_render(const qcs_params *p, const float *loc, int n, float *img, float *zimg)
{
    const int srw = p->SRImageWidth, srh = p->SRImageHigh;
    const float zoom = p->PixelZoom, ps = p->PixelSize;
    const int mode = p->RenderingMode;
    const int is3d = (p->LocType == QCS_LOCTYPE_AS3D || p->LocType == QCS_LOCTYPE_DH3D);
    const float zth = 5.0f * fmaxf(fabsf(p->MinZDepth), fabsf(p->MaxZDepth));
    float psf[49];
    for (int i = 0; i < n; i++) {
        const float *r = loc + (long)i * QCS_REC_FLOATS;
        float pp = r[QCS_POS_PPHO];
        float rx = r[QCS_POS_XPOS] * zoom, ry = r[QCS_POS_YPOS] * zoom;
        float snr = r[QCS_POS_PSNR];
        float lp = is3d ? (r[QCS_POS_CRBX] + r[QCS_POS_CRBY]) / 2.0f : r[QCS_POS_CRBX];
        float z = r[QCS_POS_ZPOS];
        if (lp != lp) { snr = 0.0f; pp = 0.0f; }
        if (lp <= 0.0f) pp = 0.0f;
        int inside = snr >= p->SNR_th && pp > 1.0f && rx > 7.0f && rx < (float)(srw - 1 - 7) &&
                     ry > 7.0f && ry < (float)(srh - 1 - 7);
        if (is3d) inside = inside && fabsf(z) < zth;
        if (!inside) continue;
        if (mode == QCS_RENDER_1PHOTON_CRLB || mode == QCS_RENDER_1PHOTON_FIXED) pp = 1.0f;
        if (mode == QCS_RENDER_1PHOTON_FIXED) lp = p->FixedLocPrec;
        if (!is3d && lp > 30.0f) { lp = 30.0f; pp = 0.0f; }
        lp = fmaxf(lp, (ps / zoom) / 2.0f);
        float rs = lp / (ps / zoom);
        render_psf(psf, rx, ry, rs);
        int ix = (int)rx, iy = (int)ry;
        if (!is3d && mode == QCS_RENDER_PEAKPHOTON_1PIXEL) {
            img[(long)srw * iy + ix] += pp;
            continue;
        }
        for (int a = 0; a < 7; a++) {
            long off = (long)srw * (iy - 3 + a) + ix - 3;
            for (int b = 0; b < 7; b++) {
                float v = pp * psf[a * 7 + b];
                img[off + b] += v;
                if (is3d && zimg) zimg[off + b] += (z * pp) * psf[a * 7 + b];
            }
        }
    }
}
