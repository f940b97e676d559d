/*
 * orc_extract.c -- ROI extraction restated for the CPU oracle (TEST INFRASTRUCTURE).
 *
 * Reference (DLL = MLEbfgs/bfgs_MLE_dll x64/bfgsMLE_dll/bfgsMLE_dll/SubregionExtraction):
 *   LDROIExtraction_ImageFiltering.cu:22-81    FilterInit (5-tap Gaussian)
 *   LDROIExtraction_ImageFiltering.cu:85-121   ImageFiltering order of passes
 *   LDROIExtraction_ImageFiltering.cu:125-282  gpuRemoveBackgroundIntensity
 *   LDROIExtraction_ImageFiltering.cu:286-470  separable convolution H / V
 *   LDROIExtraction_ImageFiltering.cu:503-544  variance sample region
 *   LDROIExtraction.cu:140-177                 ImageVarianceCalc
 *   LDROIExtraction_ROIExtraction.cu:25-177    gpuROIFindingLD
 *   LDROIExtraction_ROIExtraction.cu:181-253   gpuDetectedROIClasify
 *   LDROIExtraction_ROIExtraction.cu:257-316   gpuSubregionExtractionLD
 *   LDROIExtraction.cu:30-138                  ExtractMolecules / ROIExtraction
 *
 * Effective behaviours reproduced on purpose:
 *   - the background estimate of a pixel is averaged with its left/right neighbours only
 *     inside its aligned 32-column group (the reference's 32-thread blocks);
 *   - all clamping uses the batched (stacked) image, so filters bleed across frame seams
 *     when more than one frame is batched;
 *   - in 2D multi-emitter mode the sharpen kernel writes its own scratch input
 *     (ImageSharpen launches with (d_RawImg, d_RawImg), :675), so detection runs on the
 *     background-removed, unsmoothed image;
 *   - the variance uses exact integer sums, rounded once (the reference accumulates with
 *     float atomics in nondeterministic order).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "orc.h"

static int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

/* FilterInit (LDROIExtraction_ImageFiltering.cu:22-81).  Host FP32 without contraction;
 * expf is evaluated in double and rounded once. */
void orc_filter_taps(int R, int loc_type, int multi, float taps[5])
{
    float s7 = 0.5f, s11 = 1.0f;
    float ss = s7 + ((float)(R - 7) * (s11 - s7)) / (float)(11 - 7);
    if (ss < 0.5f) ss = 0.5f;
    if (loc_type == QCS_LOCTYPE_AS3D) ss = ss * 1.5f;
    else if (multi == 0 && loc_type == QCS_LOCTYPE_GS2D) ss = ss * 1.2f;
    ss = -1.0f / ((2.0f * ss) * ss);
    const float ctr = 2.0f;
    for (int c = 0; c < 5; c++) {
        float d = ((float)c - ctr) * ((float)c - ctr);
        float arg = d * ss;
        taps[c] = (float)exp((double)arg);
    }
    float sum = 0.0f;
    for (int i = 0; i < 5; i++)
        for (int j = 0; j < 5; j++) {
            float pr = taps[i] * taps[j];
            sum = sum + pr;
        }
    sum = sqrtf(1.0f / sum);
    for (int c = 0; c < 5; c++) taps[c] = taps[c] * sum;
}

/* Per-pixel background estimate before neighbour averaging: the mean over 8 positions of the
 * minimum of four border samples of the 9x9 window (:213-237). */
static float bkg_raw(const uint16_t *in, int w, int bh, int x, int y)
{
    float l1[8], l2[8], l3[8], l4[8];
    for (int i = 0; i < 8; i++) {
        /* top row, columns x-4+i; bottom row, columns x-3+i; right column, rows y-4+i;
         * left column, rows y-3+i -- all clamped to the batched image */
        l1[i] = in[clampi(y - 4, 0, bh - 1) * w + clampi(x - 4 + i, 0, w - 1)];
        l3[i] = in[clampi(y + 4, 0, bh - 1) * w + clampi(x - 3 + i, 0, w - 1)];
        l2[i] = in[clampi(y - 4 + i, 0, bh - 1) * w + clampi(x + 4, 0, w - 1)];
        l4[i] = in[clampi(y - 3 + i, 0, bh - 1) * w + clampi(x - 4, 0, w - 1)];
    }
    for (int i = 0; i < 8; i++) {
        l1[i] = fminf(l1[i], l2[i]);
        l3[i] = fminf(l3[i], l4[i]);
    }
    float s = 0.0f;
    for (int i = 0; i < 8; i++) s = s + fminf(l1[i], l3[7 - i]);
    return s / 8.0f;
}

/* The clamp of the x coordinate to [0, w-1] also applies to threads past the right edge of a
 * width that is not a multiple of 32: their (clamped) estimates still enter the neighbour
 * average of the last real column. */
void orc_remove_background(const uint16_t *in, uint16_t *out, int w, int bh)
{
    float *b = (float *)malloc(sizeof(float) * (size_t)(((w + 31) / 32) * 32));
    const int wt = ((w + 31) / 32) * 32;
    for (int y = 0; y < bh; y++) {
        for (int x = 0; x < wt; x++) b[x] = bkg_raw(in, w, bh, x, y);
        for (int x = 0; x < w; x++) {
            int g0 = x & ~31;
            int s0 = x - 1 < g0 ? g0 : x - 1;
            int s2 = x + 1 > g0 + 31 ? g0 + 31 : x + 1;
            float mb = ((b[s0] + b[x]) + b[s2]) / 3.0f;
            float v = (float)in[y * w + x] - mb;
            v = fmaxf(v, 0.0f);
            out[y * w + x] = (uint16_t)v;
        }
    }
    free(b);
}

/* ImageVarianceCalc + gpuImageVarianceCalcOnSelectRegion (LDROIExtraction.cu:140-177,
 * ..._ImageFiltering.cu:503-533) with exact integer sums rounded once. */
float orc_image_variance(const uint16_t *img, int w, int bh)
{
    int cx = w / 2, cy = bh / 2;
    int sw = w / 2 - 5, sh = bh / 2 - 5;
    if (sw > 256) sw = 256;
    if (sh > 256) sh = 256;
    if (sw < 1) sw = 1;
    if (sh < 1) sh = 1;
    int S = sw < sh ? sw : sh;
    unsigned long long s1 = 0, s2 = 0;
    for (int y = 0; y < S; y++)
        for (int x = 0; x < S; x++) {
            int X = cx + x, Y = cy + y;
            unsigned d1 = img[Y * w + X];
            unsigned d2 = img[Y * w + (X - S)];
            unsigned d3 = img[(Y - S) * w + X];
            unsigned d4 = img[(Y - S) * w + (X - S)];
            d1 = d1 < d2 ? d1 : d2;
            d3 = d3 < d4 ? d3 : d4;
            d1 = d1 < d3 ? d1 : d3;
            s1 += d1;
            s2 += (unsigned long long)d1 * d1;
        }
    double n = (double)S * (double)S;
    double m1 = (double)s1 / n, m2 = (double)s2 / n;
    float var = (float)(m2 - m1 * m1);
    if (var < 100.0f) var = 100.0f;
    return var;
}

/* gpuImageConv_Seperable_H (:286-359): clamp-to-edge, sum of tap * pixel from left to right,
 * truncated to uint16. */
void orc_conv_h(const uint16_t *in, uint16_t *out, int w, int bh, const float taps[5])
{
    for (int y = 0; y < bh; y++)
        for (int x = 0; x < w; x++) {
            float acc = 0.0f;
            for (int k = 0; k < 5; k++) acc = fmaf((float)in[y * w + clampi(x - 2 + k, 0, w - 1)], taps[k], acc);
            if (acc > 65535.0f) acc = 65535.0f;
            out[y * w + x] = (uint16_t)acc;
        }
}

/* gpuImageConv_Seperable_V (:382-454) */
void orc_conv_v(const uint16_t *in, uint16_t *out, int w, int bh, const float taps[5])
{
    for (int y = 0; y < bh; y++)
        for (int x = 0; x < w; x++) {
            float acc = 0.0f;
            for (int k = 0; k < 5; k++) acc = fmaf((float)in[clampi(y - 2 + k, 0, bh - 1) * w + x], taps[k], acc);
            if (acc > 65535.0f) acc = 65535.0f;
            out[y * w + x] = (uint16_t)acc;
        }
}

/* gpuROIFindingLD (LDROIExtraction_ROIExtraction.cu:25-164).  mark[] = 1 where the 3x3 test
 * passes with >= 3 of 4 diagonals; a candidate is listed when all 4 diagonals pass.
 * Candidates are listed in raster order of the batched image. */
int orc_find_rois(const uint16_t *f, int w, int bh, int h, float variance, int loc_type, int multi,
                  int R, uint8_t *mark, int *xy, int cap)
{
    float sd = sqrtf(variance);
    float th = 9.5f * sd, th1 = 3.8f * sd;
    if (loc_type == QCS_LOCTYPE_AS3D || (multi == 0 && loc_type == QCS_LOCTYPE_GS2D)) {
        th = 8.5f * sd;
        th1 = 3.5f * sd;
    }
    const int half = R / 2;
    int n = 0;
    memset(mark, 0, (size_t)w * bh);
    for (int y = 0; y < bh; y++) {
        int ry = y % h;
        for (int x = 0; x < w; x++) {
            if (!(x >= half && ry >= half && x < w - half && ry < h - half)) continue;
            int c = f[y * w + x];
            if (!((float)c >= th)) continue;
            int p[3][3];
            for (int r = 0; r < 3; r++)
                for (int k = 0; k < 3; k++) p[r][k] = f[(y - 1 + r) * w + (x - 1 + k)];
            int s8 = ((float)p[0][0] > th1) + ((float)p[0][1] > th1) + ((float)p[0][2] > th1) +
                     ((float)p[1][0] > th1) + ((float)p[1][2] > th1) +
                     ((float)p[2][0] > th1) + ((float)p[2][1] > th1) + ((float)p[2][2] > th1);
            int si = (c > p[0][1]) + (c >= p[1][2]) + (c >= p[2][1]) + (c > p[1][0]);
            int so = (c > p[0][0]) + (c > p[0][2]) + (c > p[2][0]) + (c > p[2][2]);
            if (s8 >= 7 && si == 4 && so >= 3) {
                mark[y * w + x] = 1;
                if (so == 4) {
                    if (n < cap) { xy[2 * n] = x; xy[2 * n + 1] = y; }
                    n++;
                }
            }
        }
    }
    return n;
}

/* gpuDetectedROIClasify (:181-243): multi when any mark lies in the R x R window outside the
 * central 3 x 3. */
void orc_classify_rois(const uint8_t *mark, const int *xy, int n, int multi, int R, int w, int *ft)
{
    const int half = R / 2;
    for (int i = 0; i < n; i++) {
        if (!multi) { ft[i] = QCS_ROITYPE_SINGLE; continue; }
        int x = xy[2 * i], y = xy[2 * i + 1];
        int all = 0, inner = 0;
        for (int r = 0; r < R; r++)
            for (int c = 0; c < R; c++) all += mark[(y + r - half) * w + (x + c - half)];
        for (int r = 0; r < 3; r++)
            for (int c = 0; c < 3; c++) inner += mark[(y + r - 1) * w + (x + c - 1)];
        ft[i] = (all - inner > 0) ? QCS_ROITYPE_MULTI : QCS_ROITYPE_SINGLE;
    }
}

/* gpuSubregionExtractionLD (:257-302): raw R x R pixels + header; unused header words are 0. */
void orc_cut_rois(const uint16_t *raw, const int *xy, const int *ft, int n, int R, int w, int bh, int h,
                  int start_frame, uint16_t *rois)
{
    const int W = QCS_ROI_WORDS(R), half = R / 2;
    for (int i = 0; i < n; i++) {
        uint16_t *o = rois + (long)i * W;
        int x = xy[2 * i], y = xy[2 * i + 1];
        memset(o, 0, sizeof(uint16_t) * (size_t)W);
        if (!(x >= half && y >= half && x < w - half && y < bh - half)) continue;
        for (int r = 0; r < R; r++)
            for (int c = 0; c < R; c++) o[r * R + c] = raw[(y - half + r) * w + (x - half + c)];
        int fr = start_frame + y / h;
        o[R * R + 0] = (uint16_t)x;
        o[R * R + 1] = (uint16_t)(y % h);
        o[R * R + 2] = (uint16_t)(fr % 65536);
        o[R * R + 3] = (uint16_t)(fr / 65536);
        o[R * R + 4] = (uint16_t)ft[i];
    }
}

int orc_extract(const qcs_params *p, const uint16_t *frames, int nframes, int start_frame,
                uint16_t *rois, float *wle, int cap, int *n_found, uint16_t *filtered, float *variance_out)
{
    const int w = p->ImageWidth, h = p->ImageHigh, R = p->ROISize;
    const int B = p->BatchedImageNum > 0 ? p->BatchedImageNum : 1;
    const int batch_cap = p->MaxROIPerBatch > 0 ? p->MaxROIPerBatch : QCS_MAX_POINT_NUM;
    const int W = QCS_ROI_WORDS(R);
    float taps[5];
    orc_filter_taps(R, p->LocType, p->MultiEmitterFitEn, taps);
    const size_t maxpix = (size_t)w * h * B;
    uint16_t *t0 = (uint16_t *)malloc(maxpix * 2), *t1 = (uint16_t *)malloc(maxpix * 2), *t2 = (uint16_t *)malloc(maxpix * 2);
    uint8_t *mark = (uint8_t *)malloc(maxpix);
    int *xy = (int *)malloc(sizeof(int) * 2 * (size_t)batch_cap);
    int *ft = (int *)malloc(sizeof(int) * (size_t)batch_cap);
    int total = 0, bi = 0, found_all = 0;
    for (int f0 = 0; f0 < nframes; f0 += B, bi++) {
        int nb = nframes - f0 < B ? nframes - f0 : B;
        int bh = nb * h;
        const uint16_t *raw = frames + (size_t)f0 * w * h;
        orc_remove_background(raw, t0, w, bh);
        float var = orc_image_variance(t0, w, bh);
        if (variance_out) variance_out[bi] = var;
        const uint16_t *det;
        if (p->LocType == QCS_LOCTYPE_GS2D && p->MultiEmitterFitEn > 0) {
            det = t0;   /* the sharpened image never reaches the detector */
        } else {
            orc_conv_h(t0, t1, w, bh, taps);
            orc_conv_v(t1, t2, w, bh, taps);
            det = t2;
        }
        if (filtered) memcpy(filtered, det, (size_t)w * bh * 2);
        int found = orc_find_rois(det, w, bh, h, var, p->LocType, p->MultiEmitterFitEn, R, mark, xy, batch_cap);
        int nb_rois = found < batch_cap ? found : batch_cap;
        found_all += found;
        orc_classify_rois(mark, xy, nb_rois, p->MultiEmitterFitEn, R, w, ft);
        int room = cap - total;
        if (room < 0) room = 0;
        int keep = nb_rois < room ? nb_rois : room;
        if (keep > 0) {
            orc_cut_rois(raw, xy, ft, keep, R, w, bh, h, start_frame + f0, rois + (size_t)total * W);
            if (wle) orc_wle(rois + (size_t)total * W, keep, R, p->LocType, wle + (size_t)total * QCS_WLE_FLOATS);
        }
        total += keep;
    }
    if (n_found) *n_found = found_all;
    free(t0); free(t1); free(t2); free(mark); free(xy); free(ft);
    return total;
}
