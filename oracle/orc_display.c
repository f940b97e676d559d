/*
 * orc_display.c -- the colour-coded 3D super-resolution image, restated for the CPU oracle
 * (TEST INFRASTRUCTURE).
 *
 * Reference (DLL = MLEbfgs/bfgs_MLE_dll x64/bfgsMLE_dll/bfgsMLE_dll/ImgRender):
 *   DLL/ImageRender_3d.cu:238-346    RenderSaveImg_3D (per-pixel colour, encodings)
 *   DLL/ImageRender_3d.cu:375-403    ImgRender_GetSaveImg_3D (brightness scale)
 *   DLL/ImageRender_3d.cu:424-455    ColorGenerate (z -> hue, intensity -> value)
 *   DLL/ImageRender_3d.cu:458-520    HSVtoRGB
 *   DLL/ImageRender_3d.cu:529-590    gpuGetSRImageHistgram, GetSRImageHistgram (97% level)
 *   DLL/ImageRender_base.h:43-57     encodings, colour modes;  ImageRender_top.h:31 2048 bins
 * Histogram counts are integers here (the reference counts with float atomics, which agree while
 * a bin holds fewer than 2^24 pixels); the 97% search sums them in float as the reference does.
 */
#include <limits.h>
#include <stdint.h>
#include "orc.h"

#define HIST_N   2048
#define HIST_MAX 12000.0f

/* float -> int with saturation and NaN -> 0 (the device conversion) */
static int f2i_sat(float x)
{
    if (x != x) return 0;
    if (x >= 2147483648.0f) return INT_MAX;
    if (x <= -2147483648.0f) return INT_MIN;
    return (int)x;
}

/* GetSRImageHistgram: the intensity below which 97% of the histogrammed pixels lie */
float orc_display_max_intensity(const float *img, long long npx, const int *hist_in, int *hist_out)
{
    int h[HIST_N];
    const float gap = HIST_MAX / HIST_N;
    for (int k = 0; k < HIST_N; k++) h[k] = 0;
    if (hist_in) {
        for (int k = 0; k < HIST_N; k++) h[k] = hist_in[k];
    } else {
        for (long long i = 0; i < npx; i++) {
            const int pos = f2i_sat(img[i] / gap - 1);
            if (pos >= 0 && pos < HIST_N) h[pos]++;
        }
    }
    if (hist_out)
        for (int k = 0; k < HIST_N; k++) hist_out[k] = h[k];
    float tot = 0, cur = 0;
    float gapi = (float)(HIST_N - 1);
    for (int k = 0; k < HIST_N; k++) tot += (float)h[k];
    for (int k = 0; k < HIST_N; k++) {
        cur += (float)h[k];
        if (cur / tot > 0.97f) {
            gapi = (float)k;
            break;
        }
    }
    return gapi * gap;
}

static void hsv_to_rgb(int h, int s, int v, int *r, int *g, int *b)
{
    if (h < 0) h = 0;
    if (s < 0) s = 0;
    if (v < 0) v = 0;
    const float mx = (float)v * 2.55f;
    const float mn = mx * (float)(100 - s) / 100.0f;
    const int i = h / 60, d = h % 60;
    const float adj = (mx - mn) * (float)d / 60.0f;
    switch (i) {
    case 0: *r = (int)(mx + 0.5f); *g = (int)(mn + adj + 0.5f); *b = (int)(mn + 0.5f); break;
    case 1: *r = (int)(mx - adj + 0.5f); *g = (int)(mx + 0.5f); *b = (int)(mn + 0.5f); break;
    case 2: *r = (int)(mn + 0.5f); *g = (int)(mx + 0.5f); *b = (int)(mn + adj + 0.5f); break;
    case 3: *r = (int)(mn + 0.5f); *g = (int)(mx - adj + 0.5f); *b = (int)(mx + 0.5f); break;
    case 4: *r = (int)(mn + adj + 0.5f); *g = (int)(mn + 0.5f); *b = (int)(mx + 0.5f); break;
    default: *r = (int)(mx + 0.5f); *g = (int)(mn + 0.5f); *b = (int)(mx - adj + 0.5f); break;
    }
    *r = *r < 0 ? 0 : (*r > 255 ? 255 : *r);
    *g = *g < 0 ? 0 : (*g > 255 ? 255 : *g);
    *b = *b < 0 ? 0 : (*b > 255 ? 255 : *b);
}

/* RenderSaveImg_3D with ColorGenerate; mode 0 BGR, 1 RGB (3 bytes), 2 BGRA, 3 RGBA (4 bytes) */
void orc_display_rgb(const qcs_params *p, const float *img, const float *zimg, long long npx, float max_intensity, int mode,
                     uint8_t *out)
{
    const int nb = mode < 2 ? 3 : 4;
    const float zr = p->MaxZDepth - p->MinZDepth;
    for (long long i = 0; i < npx; i++) {
        uint8_t *o = out + i * nb;
        const float it = img[i];
        if (it == 0) {
            o[0] = o[1] = o[2] = 0;
            if (nb == 4) o[3] = 0xff;
            continue;
        }
        float z = zimg[i] / img[i];
        int v = f2i_sat(it * 120.0f / max_intensity);
        if (v >= 100) v = 100;
        if (z < p->MinZDepth) z = p->MinZDepth;
        if (z > p->MaxZDepth) z = p->MaxZDepth;
        int h = f2i_sat((z - p->MinZDepth) * 240.0f / zr);
        if (p->ColorMode_3D == 0) h = 240 - h;
        if (h < 0) h = 0;
        if (h > 240) h = 240;
        int r, g, b;
        hsv_to_rgb(h, 100, v, &r, &g, &b);
        if (mode == 0 || mode == 2) {
            o[0] = (uint8_t)b; o[1] = (uint8_t)g; o[2] = (uint8_t)r;
        } else {
            o[0] = (uint8_t)r; o[1] = (uint8_t)g; o[2] = (uint8_t)b;
        }
        if (nb == 4) o[3] = 0xff;
    }
}
