/*
 * orc.h -- CPU oracle for the QC-STORM localization hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker the tests, the smoke check and the
 * benchmark's CPU-baseline leg compare the HIP library against.  The product library
 * (libqcstorm.so) never links, loads or calls it.
 *
 * It is a plain-C restatement of the reference algorithm, function by function (each function
 * cites the reference file:line it follows).  Paths are relative to
 *   DLL = MLEbfgs/bfgs_MLE_dll x64/bfgsMLE_dll/bfgsMLE_dll
 * Arithmetic: FP32 with every multiply-add written as an explicit fmaf (built with
 * -ffp-contract=off), IEEE division and sqrt.  Transcendentals: the reference's __expf/__logf
 * cannot be reproduced; the oracle offers
 *   ORC_TMODE_DET  -- orc_qexpf/orc_qlogf (and the fit loss's table log orc_tlog_u),
 *                     deterministic FP32 routines (<= 2 ulp, pinned against
 *                     double-precision exp/log by the tests);
 *   ORC_TMODE_LIBM -- the C library's expf/logf.
 * and two evaluations of the Gaussian model inside the fit loss:
 *   ORC_MODEL_SEPARABLE -- exp(-s dx^2) * exp(-s dy^2) (one exp per row/column);
 *   ORC_MODEL_LITERAL   -- exp(-s (dx^2 + dy^2)) per pixel, as written in the reference.
 * Both are the same function up to FP32 rounding.  Every fitter also exists in FP64
 * (the same schedule in double precision) to measure FP32 error.
 */
#ifndef QCS_ORACLE_H
#define QCS_ORACLE_H

#include <stdint.h>
#include "../include/qcstorm.h"

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_TMODE_DET        0
#define ORC_TMODE_LIBM       1
#define ORC_MODEL_SEPARABLE  0
#define ORC_MODEL_LITERAL    1

void  orc_set_modes(int tmode, int model);
void  orc_get_modes(int *tmode, int *model);

/* deterministic FP32 exp / natural log */
float orc_qexpf(float x);
float orc_qlogf(float x);
void  orc_qexpf_v(const float *x, float *y, long n);
void  orc_qlogf_v(const float *x, float *y, long n);
/* table-driven log of the fit loss: orc_tlog_u(x, c) = log x + c (c added before the final
 * e ln2 step, as the HIP loss does); orc_tlogf(x) = orc_tlog_u(x, 0); the 32 table pairs */
float orc_tlog_u(float x, float c);
float orc_tlogf(float x);
void  orc_tlogf_v(const float *x, float *y, long n);
void  orc_tlog_table(uint32_t out[64]);
/* binade table log of the one-emitter losses: orc_tlogb_u(x, c) = log x + c; the 4096 pairs */
float orc_tlogb_u(float x, float c);
float orc_tlogbf(float x);
void  orc_tlogbf_v(const float *x, float *y, long n);
void  orc_tlogb_table(uint32_t out[8192]);
/* the CRLB's narrow-range division by dh = 0.02 and a count of its differences from x / 0.02f */
float orc_div_dh_fast(float x);
long long orc_div_dh_scan(uint32_t b0, uint32_t b1);

/* ---------------- extraction (DLL/SubregionExtraction) ---------------- */
void  orc_filter_taps(int roi_size, int loc_type, int multi, float taps[5]);
void  orc_remove_background(const uint16_t *in, uint16_t *out, int w, int bh);
float orc_image_variance(const uint16_t *img, int w, int bh);
void  orc_conv_h(const uint16_t *in, uint16_t *out, int w, int bh, const float taps[5]);
void  orc_conv_v(const uint16_t *in, uint16_t *out, int w, int bh, const float taps[5]);
/* returns the number of candidates found; writes min(found, cap) (x, y) pairs in raster order
 * and the mark image (uint8, w*bh). */
int   orc_find_rois(const uint16_t *filt, int w, int bh, int h, float variance, int loc_type,
                    int multi, int roi_size, uint8_t *mark, int *xy, int cap);
void  orc_classify_rois(const uint8_t *mark, const int *xy, int n, int multi, int roi_size, int w,
                        int *fit_type);
void  orc_cut_rois(const uint16_t *raw, const int *xy, const int *fit_type, int n, int roi_size,
                   int w, int bh, int h, int start_frame, uint16_t *rois);
/* Whole extraction of nframes frames in batches of p->BatchedImageNum, followed by the WLE
 * pre-estimation of each batch (the reference always runs it: WLE_ENABLE 1).  Each batch keeps
 * at most p->MaxROIPerBatch candidates.  Returns the number of ROI records written (<= cap);
 * *n_found (may be NULL) receives the number of candidates found.  `filtered` (may be NULL)
 * receives the detection image of the last batch, `variance_out` (may be NULL) the per-batch
 * variances. */
int   orc_extract(const qcs_params *p, const uint16_t *frames, int nframes, int start_frame,
                  uint16_t *rois, float *wle, int cap, int *n_found, uint16_t *filtered,
                  float *variance_out);

/* mixed-size extraction of the converged multi-emitter path (orc_mixed.c): detection at size 7,
 * the library's size rule per candidate, the first cap[s] candidates of each size cut into
 * rois[s] (sizes 7, 9, 11); candidates examined: the first cand_cap of the call */
int   orc_mixed_size(const uint8_t *mark, int w, int h, int x, int y, const float *wle7);
int   orc_extract_mixed(const qcs_params *p, const uint16_t *frames, int nframes, int start_frame, int cand_cap,
                        uint16_t *const rois[3], const int cap[3], int counts[3], int *sizes, int *n_cand,
                        int *n_found);

/* ---------------- WLE pre-estimation (DLL/WLEParaEstimation) ---------------- */
void  orc_spline_interp(const float *x_i, const float *y_i, int n_in, const float *x_o, float *y_o,
                        int n_out);
void  orc_wle_psf_width(const uint16_t *rois, int n, int roi_size, float *wle);
void  orc_wle_nearest(const uint16_t *rois, int n, int roi_size, float *wle);
void  orc_wle_classify(const uint16_t *rois, int n, int roi_size, int loc_type, float *wle);
void  orc_wle(const uint16_t *rois, int n, int roi_size, int loc_type, float *wle);

/* ---------------- fitting (DLL/bfgsMLE_Loc_LD, DLL/bfgsMLE_Loc_HD, DLL/bfgsMLE_Loc_top) ---- */
typedef struct orc_fit_stats {
    int n_in, n_single, n_2e, n_3e, n_rejected, n_rows;
} orc_fit_stats;
/* full cascade + CRLB; returns rows written (oValidFluoNum) */
int   orc_fit(const qcs_params *p, const uint16_t *rois, const float *wle, int n, float *loc,
              int loc_cap, orc_fit_stats *st);
/* single-emitter mode, OpenMP over ROIs (CPU baseline); -1 in multi-emitter mode */
int   orc_fit_omp(const qcs_params *p, const uint16_t *rois, const float *wle, int n, float *loc, int threads);
/* same schedule in FP64 (transcendentals from libm, double); loc in double */
int   orc_fit_f64(const qcs_params *p, const uint16_t *rois, const float *wle, int n, double *loc,
                  int loc_cap, orc_fit_stats *st);
/* single-kernel entry points: fit the listed ROIs with one model and return the raw fitted
 * parameter vectors (scaled units, as Ininf at the end of the BFGS loop); `iters` overrides
 * the reference iteration count when > 0.  model: 0 GS2D-1E, 1 AS3D-1E, 2 GS2D-2E, 3 GS2D-3E,
 * 4 AS3D-2E.  Returns the parameter count. */
int   orc_fit_params(const qcs_params *p, int model, const uint16_t *rois, const float *wle, int n,
                     float *theta, int iters);
int   orc_fit_params_f64(const qcs_params *p, int model, const uint16_t *rois, const float *wle,
                         int n, double *theta, int iters);
/* loss value of `model` at theta for ROI 0 (scaled units); FP32 and FP64 */
float orc_loss(const qcs_params *p, int model, const uint16_t *roi, const float *wle, const float *theta);
/* loss evaluations made by the calling thread since the last reset */
long  orc_loss_evals(int reset);
/* acceptance of a one-emitter fit vector x (returns 1 and writes the row if accepted) and the
 * emitter selection of a multi-emitter fit vector (returns 0, 1 or 2 rows written) */
int   orc_select_1e(const qcs_params *p, int model, const float *x, const uint16_t *roi, float *row);
int   orc_select_me(const qcs_params *p, int model, const float *x, float nn, const uint16_t *roi, float *row,
                    float *row2);

/* ---------------- CRLB (DLL/bfgsMLE_Loc_top/LocPrecCalc_GaussianCRLB.h) ---------------- */
void  orc_crlb(const qcs_params *p, float *loc, int n);
void  orc_crlb_f64(const qcs_params *p, double *loc, int n);

/* ---------------- compaction / render ---------------- */
int   orc_compact_sorted(const float *loc, int n, int first, int end, float *out);
int   orc_compact(const float *loc, int n, float *out);
void  orc_render(const qcs_params *p, const float *loc, int n, float *img, float *zimg);

/* ---------------- FP64 maximum-likelihood truth tools ---------------- */
/* Converged unweighted Poisson MLE of K emitters (K = 1..3) sharing sigma (2D) or with
 * per-emitter sigma_x/sigma_y (astig = 1) and one background, in photon units, pixel-edge
 * coordinates.  theta layout: [N_k, x_k, y_k, (sx_k, sy_k | sigma shared)]..., bg.
 * Returns iterations used (<0 on failure).  loglik receives the final log-likelihood. */
int   orc_mle_f64(const double *img, int R, int K, int astig, double *theta, int max_iter,
                  double *loglik);
/* Full-Fisher CRLB (standard deviations) at theta for the same models. */
int   orc_fisher_crlb_f64(int R, int K, int astig, const double *theta, double *crlb);

/* statistics of a batch of rows (orc_stats.c) */
int   orc_loc_stats(const qcs_params *p, const float *loc, int n, qcs_loc_stats *out);
float orc_activation_density(float r0, float r1, float radius_um);
float orc_hist_mean(const int *h, int len, float pct);

/* consecutive-frame merging (orc_consec.c) */
typedef struct orc_consec orc_consec;
void  orc_consec_links(const float *loc, int n, float th_px, int *back, int *fwd);
orc_consec *orc_consec_create(void);
void  orc_consec_destroy(orc_consec *c);
int   orc_consec_held(const orc_consec *c);
int   orc_consec_fit(const qcs_params *p, orc_consec *c, const float *rows, int n, int is_end, float *out, int out_cap,
                     float *ontime_ratio);

/* colour-coded 3D image (orc_display.c) */
float  orc_display_max_intensity(const float *img, long long npx, const int *hist_in, int *hist_out);
void   orc_display_rgb(const qcs_params *p, const float *img, const float *zimg, long long npx, float max_intensity, int mode,
                       uint8_t *out);

/* drift correction by SR-image cross-correlation (orc_drift.c) */
int    orc_drift_total_frame(const float *loc, long long n);
int    orc_drift_groups(int total, int corr, int *gs, int *ge);
void   orc_drift_group_rows(const float *loc, long long n, int ng, const int *gs, const int *ge, long long *first,
                            long long *last);
void   orc_drift_render(const qcs_params *p, const float *loc, long long n, int srw, int srh, float *img);
double orc_drift_cc(const float *img1, const float *img2, int w, int h, int sx, int sy);
void   orc_drift_peak(const double *cc, int bias_x, int bias_y, float *shift_x, float *shift_y);
void   orc_drift_interp(int total, int ng, const int *gs, const int *ge, const float *gx, const float *gy, float *fx, float *fy);
int    orc_drift_correct(const qcs_params *p, float *loc, long long n, int corr, float *fx, float *fy, int fcap,
                         int *total_out, float *gx_out, float *gy_out);

#ifdef __cplusplus
}
#endif
#endif
