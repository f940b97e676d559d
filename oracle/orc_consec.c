/*
 * orc_consec.c -- consecutive-frame merging of localization rows, restated for the CPU oracle
 * (TEST INFRASTRUCTURE).
 *
 * Reference (DLL = MLEbfgs/bfgs_MLE_dll x64/bfgsMLE_dll/bfgsMLE_dll):
 *   DLL/ConsecutiveFit/ConsecutiveFit_top.cu:21-101   ConsecutiveFit_WeightedAvg (held-back group,
 *                                                     on-time ratios, CRLB of the output rows)
 *   DLL/ConsecutiveFit/ConsecutiveFit.cu:18-77        gpuFindConsecutiveMolecules
 *   DLL/ConsecutiveFit/ConsecutiveFit.cu:81-188       gpuConsecutiveFit
 *   DLL/ConsecutiveFit/ConsecutiveFit.cu:191-220      gpuRemoveConsecutiveFluo
 *   DLL/ConsecutiveFit/ConsecutiveFit.h:31            ThreadsPerBlock = 32 (start of the scan)
 *   DLL/StatisticsInfo/statisticsInfo.h:65            MaxOnTimeConsecutiveNum = 10
 * The link search is the reference's loop, literally: every row scans from the first row of its
 * 32-row block while the frame read stays <= its frame + 1, on integer-truncated positions.
 * The weighted sums use fmaf where the reference build contracts a multiply-add (nvcc's default
 * -fmad=true on `w += x * p`).
 */
#include <limits.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "orc.h"

#define CONSEC_BLOCK 32
#define ONTIME_BINS  10

/* float -> int with the saturating, NaN -> 0 semantics of the GPU conversion */
static int f2i(float x)
{
    if (x != x) return 0;
    if (x >= 2147483648.0f) return INT_MAX;
    if (x <= -2147483648.0f) return INT_MIN;
    return (int)x;
}

/* squared integer distance as the reference computes it: int products (wrapping), then float */
static float idist2(int x1, int y1, int x2, int y2)
{
    const unsigned dx = (unsigned)x1 - (unsigned)x2, dy = (unsigned)y1 - (unsigned)y2;
    return (float)(int)(dx * dx + dy * dy);
}

/* gpuFindConsecutiveMolecules (ConsecutiveFit.cu:18-77): back[i] = closest row of the next frame
 * (first in row order among equal distances) or -2; fwd[j] = 1 for every row some row links to */
void orc_consec_links(const float *loc, int n, float th_px, int *back, int *fwd)
{
    memset(back, 0, sizeof(int) * (size_t)n);
    memset(fwd, 0, sizeof(int) * (size_t)n);
    for (int i = 0; i < n; i++) {
        const float *r = loc + (size_t)i * QCS_REC_FLOATS;
        const int x1 = f2i(r[QCS_POS_XPOS]), y1 = f2i(r[QCS_POS_YPOS]), f1 = f2i(r[QCS_POS_FRME]);
        int f2 = 0, paired = -1, found = 0;
        float th2 = th_px;
        th2 = th2 * th2;
        for (int j = (i / CONSEC_BLOCK) * CONSEC_BLOCK; j < n && f2 <= f1 + 1; j++) {
            const float *q = loc + (size_t)j * QCS_REC_FLOATS;
            const int x2 = f2i(q[QCS_POS_XPOS]), y2 = f2i(q[QCS_POS_YPOS]);
            f2 = f2i(q[QCS_POS_FRME]);
            const float d2 = idist2(x1, y1, x2, y2);
            if (d2 < th2 && f2 - f1 == 1) {
                paired = j;
                found++;
                th2 = d2;
            }
        }
        if (found > 0) {
            back[i] = paired;
            fwd[paired] = 1;
        } else {
            back[i] = -2;
        }
    }
}

struct orc_consec {
    float *last;
    int n_last, cap;
};

orc_consec *orc_consec_create(void)
{
    return (orc_consec *)calloc(1, sizeof(orc_consec));
}

void orc_consec_destroy(orc_consec *c)
{
    if (!c) return;
    free(c->last);
    free(c);
}

int orc_consec_held(const orc_consec *c) { return c->n_last; }

/* ConsecutiveFit_WeightedAvg (ConsecutiveFit_top.cu:21-101).  out receives the rows given back by
 * this call (the held-back group of the previous call, or everything when is_end), before the
 * removal of zero rows; returns their number, or -1 when out_cap is too small. */
int orc_consec_fit(const qcs_params *p, orc_consec *c, const float *rows, int n, int is_end, float *out, int out_cap,
                   float *ontime_ratio)
{
    const int L = c->n_last, N = L + n;
    const int out_n = is_end ? N : L;
    if (out_n > out_cap) return -1;
    float *a = (float *)malloc(sizeof(float) * QCS_REC_FLOATS * (size_t)(N > 0 ? N : 1));
    int *back = (int *)malloc(sizeof(int) * (size_t)(N > 0 ? N : 1));
    int *fwd = (int *)malloc(sizeof(int) * (size_t)(N > 0 ? N : 1));
    if (L) memcpy(a, c->last, sizeof(float) * QCS_REC_FLOATS * (size_t)L);
    if (n) memcpy(a + (size_t)L * QCS_REC_FLOATS, rows, sizeof(float) * QCS_REC_FLOATS * (size_t)n);
    const float th_px = p->ConsecFit_DistanceTh_nm / p->PixelSize;
    orc_consec_links(a, N, th_px, back, fwd);

    /* gpuConsecutiveFit (ConsecutiveFit.cu:81-188): heads (no row links to them) take the
     * 10/CRLBx-weighted average of their chain; sums of photons and background */
    int hist[ONTIME_BINS] = {0}, valid = 0;
    for (int i = 0; i < N; i++) {
        if (fwd[i] != 0) continue;
        float wx = 0, wy = 0, wz = 0, wsx = 0, wsy = 0, sp = 0, st = 0, sb = 0, sw = 0;
        int cur = i, cnt = 0, next = 1;
        while (next) {
            const float *r = a + (size_t)cur * QCS_REC_FLOATS;
            const float prec = 10 / r[QCS_POS_CRBX];
            sp += r[QCS_POS_PPHO];
            st += r[QCS_POS_TPHO];
            sb += r[QCS_POS_BAKG];
            wx = fmaf(r[QCS_POS_XPOS], prec, wx);
            wy = fmaf(r[QCS_POS_YPOS], prec, wy);
            wz = fmaf(r[QCS_POS_ZPOS], prec, wz);
            wsx = fmaf(r[QCS_POS_SIGX], prec, wsx);
            wsy = fmaf(r[QCS_POS_SIGY], prec, wsy);
            sw += prec;
            cur = back[cur];
            next = cur > 0;
            cnt++;
        }
        wx /= sw;
        wy /= sw;
        wz /= sw;
        wsx /= sw;
        wsy /= sw;
        const float snr = sqrtf(p->QE) * sp / sqrtf(sp + sb);
        float *h = a + (size_t)i * QCS_REC_FLOATS;
        h[QCS_POS_PPHO] = sp;
        h[QCS_POS_XPOS] = wx;
        h[QCS_POS_YPOS] = wy;
        h[QCS_POS_ZPOS] = wz;
        h[QCS_POS_SIGX] = wsx;
        h[QCS_POS_SIGY] = wsy;
        h[QCS_POS_TPHO] = st;
        h[QCS_POS_BAKG] = sb;
        h[QCS_POS_PSNR] = snr;
        if (i >= L) {
            int on = cnt - 1;
            on = on < 0 ? 0 : (on > ONTIME_BINS - 1 ? ONTIME_BINS - 1 : on);
            valid++;
            hist[on]++;
        }
    }
    /* gpuRemoveConsecutiveFluo (ConsecutiveFit.cu:191-220): zero the rest of every chain */
    for (int i = 0; i < N; i++) {
        if (fwd[i] != 0) continue;
        for (int nx = back[i]; nx > 0; nx = back[nx]) memset(a + (size_t)nx * QCS_REC_FLOATS, 0, sizeof(float) * QCS_REC_FLOATS);
    }
    if (ontime_ratio) {
        const int v = valid <= 0 ? 1 : valid;
        for (int k = 0; k < ONTIME_BINS; k++) ontime_ratio[k] = (float)hist[k] * 1.0f / (float)v;
    }
    /* CRLB of the rows of the previous group (LocPrecCalc_GaussianCRLB on FluoNum_LastGroup rows) */
    orc_crlb(p, a, L);
    if (out_n) memcpy(out, a, sizeof(float) * QCS_REC_FLOATS * (size_t)out_n);
    if (is_end) {
        c->n_last = 0;
    } else {
        if (n > c->cap) {
            free(c->last);
            c->last = (float *)malloc(sizeof(float) * QCS_REC_FLOATS * (size_t)n);
            c->cap = n;
        }
        if (n) memcpy(c->last, a + (size_t)L * QCS_REC_FLOATS, sizeof(float) * QCS_REC_FLOATS * (size_t)n);
        c->n_last = n;
    }
    free(a);
    free(back);
    free(fwd);
    return out_n;
}
