/*
 * orc_mixed.c -- mixed-size ROI extraction of the converged multi-emitter path, restated for
 * the CPU oracle (TEST INFRASTRUCTURE).
 *
 * No reference counterpart: the reference cuts every ROI at one size per run
 * (LDROIExtraction_ROIExtraction.cu:257-302).  Detection, classification and WLE are the
 * reference's (orc_extract.c, orc_wle.c) at the base size 7; the size of each candidate's ROI
 * follows the library's rule, stated here in plain loops:
 *
 *   s    = min(SigmaL, SigmaR, SigmaU, SigmaD) of the candidate's 7 x 7 WLE record
 *   Rpsf = 7 if s <= 1.4, 9 if s <= 1.9, else 11                    (about +-2.5 sigma)
 *   c(d) = detector marks in the (2d+1)^2 window around the candidate, rows of its own frame
 *          only (c(0) = 1, the candidate itself)
 *   dmin = smallest d in 1..5 with c(d) > 1 (6 if none)
 *   Rnb  = 2 dmin + 3 when dmin is 3 or 4 (a neighbour at the edge of the 7 x 7 window is taken
 *          in whole, with a pixel of margin), else 7
 *   R    = max(Rpsf, Rnb); then while R > 7 and the R x R window holds more than 2 other marks
 *          (c(R/2) - 1 > 2: more emitters than a 3-emitter fit models), R -= 2; then while
 *          R > 7 and the window leaves the frame, R -= 2.
 *
 * The call examines its first cand_cap candidates (raster order, at most MaxROIPerBatch per
 * batch, as the extraction's capacity); each set (7, 9, 11) lists its candidates in that order,
 * at most cap_s of them; records are the reference's ROI record at that size, with
 * FitType = 1 when a mark lies in the window outside the central 3 x 3.
 */
#include <stdlib.h>
#include <string.h>
#include "orc.h"

/* marks in rows y0..y1 (inclusive, clamped to [ylo, yhi]) and columns x0..x1 (clamped to the image) */
static int count_marks(const uint8_t *mark, int w, int ylo, int yhi, int x0, int x1, int y0, int y1)
{
    int n = 0;
    y0 = y0 < ylo ? ylo : y0;
    y1 = y1 > yhi ? yhi : y1;
    x0 = x0 < 0 ? 0 : x0;
    x1 = x1 > w - 1 ? w - 1 : x1;
    for (int y = y0; y <= y1; y++)
        for (int x = x0; x <= x1; x++) n += mark[(size_t)y * w + x];
    return n;
}

int orc_mixed_size(const uint8_t *mark, int w, int h, int x, int y, const float *wle7)
{
    const int fy0 = (y / h) * h, fy1 = fy0 + h - 1, ry = y - fy0;
    float s = wle7[QCS_WLE_SIGMAL];
    s = wle7[QCS_WLE_SIGMAR] < s ? wle7[QCS_WLE_SIGMAR] : s;
    s = wle7[QCS_WLE_SIGMAU] < s ? wle7[QCS_WLE_SIGMAU] : s;
    s = wle7[QCS_WLE_SIGMAD] < s ? wle7[QCS_WLE_SIGMAD] : s;
    const int rpsf = (s <= 1.4f) ? 7 : ((s <= 1.9f) ? 9 : 11);
    int c[6];
    for (int d = 0; d <= 5; d++) c[d] = count_marks(mark, w, fy0, fy1, x - d, x + d, y - d, y + d);
    int dmin = 6;
    for (int d = 5; d >= 1; d--)
        if (c[d] > 1) dmin = d;
    const int rnb = (dmin == 3 || dmin == 4) ? 2 * dmin + 3 : 7;
    int R = rpsf > rnb ? rpsf : rnb;
    while (R > 7 && c[R / 2] - 1 > 2) R -= 2;
    while (R > 7 && !(x - R / 2 >= 0 && x + R / 2 < w && ry - R / 2 >= 0 && ry + R / 2 < h)) R -= 2;
    return R;
}

int orc_extract_mixed(const qcs_params *p0, const uint16_t *frames, int nframes, int start_frame, int cand_cap,
                      uint16_t *const rois[3], const int cap[3], int counts[3], int *sizes, int *n_cand,
                      int *n_found)
{
    qcs_params pp = *p0;
    pp.ROISize = 7;
    const qcs_params *p = &pp;
    const int w = p->ImageWidth, h = p->ImageHigh;
    const int B = p->BatchedImageNum > 0 ? p->BatchedImageNum : 1;
    const int batch_cap = p->MaxROIPerBatch > 0 ? p->MaxROIPerBatch : QCS_MAX_POINT_NUM;
    const int RS[3] = {7, 9, 11};
    float taps[5];
    orc_filter_taps(7, p->LocType, p->MultiEmitterFitEn, taps);
    const size_t maxpix = (size_t)w * h * B;
    uint16_t *t0 = (uint16_t *)malloc(maxpix * 2), *t1 = (uint16_t *)malloc(maxpix * 2), *t2 = (uint16_t *)malloc(maxpix * 2);
    uint8_t *mark = (uint8_t *)malloc(maxpix);
    int *xy = (int *)malloc(sizeof(int) * 2 * (size_t)batch_cap);
    int *ft = (int *)calloc((size_t)batch_cap, sizeof(int));
    uint16_t *r7 = (uint16_t *)malloc(sizeof(uint16_t) * QCS_ROI_WORDS(7) * (size_t)batch_cap);
    float *w7 = (float *)malloc(sizeof(float) * QCS_WLE_FLOATS * (size_t)batch_cap);
    int found_all = 0, cand = 0;
    counts[0] = counts[1] = counts[2] = 0;
    for (int f0 = 0; f0 < nframes; f0 += B) {
        const int nb = nframes - f0 < B ? nframes - f0 : B;
        const int bh = nb * h;
        const uint16_t *raw = frames + (size_t)f0 * w * h;
        orc_remove_background(raw, t0, w, bh);
        const float var = orc_image_variance(t0, w, bh);
        const uint16_t *det;
        if (p->LocType == QCS_LOCTYPE_GS2D && p->MultiEmitterFitEn > 0) {
            det = t0;
        } else {
            orc_conv_h(t0, t1, w, bh, taps);
            orc_conv_v(t1, t2, w, bh, taps);
            det = t2;
        }
        const int found = orc_find_rois(det, w, bh, h, var, p->LocType, p->MultiEmitterFitEn, 7, mark, xy, batch_cap);
        int nk = found < batch_cap ? found : batch_cap;
        if (nk > cand_cap - cand) nk = cand_cap - cand > 0 ? cand_cap - cand : 0;   /* the call's candidate capacity */
        found_all += found;
        /* WLE of the candidates' 7 x 7 cuts (nearest neighbours per frame over the batch's list) */
        orc_classify_rois(mark, xy, nk, p->MultiEmitterFitEn, 7, w, ft);
        orc_cut_rois(raw, xy, ft, nk, 7, w, bh, h, start_frame + f0, r7);
        orc_wle(r7, nk, 7, p->LocType, w7);
        for (int i = 0; i < nk; i++, cand++) {
            const int x = xy[2 * i], y = xy[2 * i + 1];
            const int R = orc_mixed_size(mark, w, h, x, y, w7 + (size_t)i * QCS_WLE_FLOATS);
            const int s = (R - 7) / 2;
            if (sizes) sizes[cand] = R;
            if (counts[s] >= cap[s]) continue;
            int ftype;
            orc_classify_rois(mark, xy + 2 * i, 1, p->MultiEmitterFitEn, R, w, &ftype);
            orc_cut_rois(raw, xy + 2 * i, &ftype, 1, R, w, bh, h, start_frame + f0,
                         rois[s] + (size_t)counts[s] * QCS_ROI_WORDS(RS[s]));
            counts[s]++;
        }
    }
    if (n_cand) *n_cand = cand;
    if (n_found) *n_found = found_all;
    free(t0); free(t1); free(t2); free(mark); free(xy); free(ft); free(r7); free(w7);
    return counts[0] + counts[1] + counts[2];
}
