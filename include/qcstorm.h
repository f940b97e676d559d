/*
 * qcstorm.h -- C-ABI of the MI355X (gfx950) localization library libqcstorm.so.
 *
 * The library replaces the localization hot path of QC-STORM: ROI extraction from
 * uint16 sCMOS frames, weighted-likelihood pre-estimation, BFGS Poisson-MLE Gaussian
 * fitting (2D / astigmatic 3D, 1-3 emitters) with CRLB, frame-sorted compaction of
 * the 12-float localization records and super-resolution splat rendering.
 *
 * Every entry point is extern "C", takes plain pointers and sizes, and returns an
 * int status (QCS_OK == 0).  qcs_last_error() gives a thread-local message for the
 * last failing call.  Pointers named d_* are device (HBM) pointers; the others are
 * host pointers.  `stream` is a hipStream_t passed as void* (NULL = default stream).
 *
 * Reference interfaces replaced (paths relative to the reference tree):
 *   DLL = MLEbfgs/bfgs_MLE_dll x64/bfgsMLE_dll/bfgsMLE_dll
 *   REL = MLEbfgs/bfgs_MLE_dll x64/bfgsMLE_dll/x64/Release
 *   IJ  = Plugin for ImageJ/QC-STORM_ CPPDLL/QC-STORM_ CPPDLL
 *   MM  = Plugin for Micro-Manager/QC-STORM CPPDLL/QC-STORM CPPDLL
 */
#ifndef QCSTORM_H
#define QCSTORM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes (the reference prints and continues; see DLL/cudaWrapper/cudaWrapper.cu:129-139) */
#define QCS_OK               0
#define QCS_ERR_ARG          1  /* invalid argument (NULL pointer, bad size, unsupported ROI size) */
#define QCS_ERR_HIP          2  /* HIP runtime error */
#define QCS_ERR_CAPACITY     3  /* output capacity too small; partial results are reported */
#define QCS_ERR_STATE        4  /* call not valid in the current handle state */
#define QCS_ERR_NOMEM        5  /* allocation failed */
#define QCS_ERR_NODEVICE     6  /* no usable gfx950 device */
#define QCS_ERR_IO           7  /* result file could not be created or written */

/* ---- localization types (DLL/bfgsMLE_Loc_LD/bfgs_CommonPara.h:274-277) */
#define QCS_LOCTYPE_GS2D     0
#define QCS_LOCTYPE_AS3D     1
#define QCS_LOCTYPE_DH3D     2  /* accepted and fitted as GS2D, as LocType_IsGS2D does */

/* ---- 12-float localization record (bfgs_CommonPara.h:292-309) */
#define QCS_REC_FLOATS       12
#define QCS_POS_PPHO         0   /* peak photon */
#define QCS_POS_XPOS         1   /* x, pixel (pixel i spans [i, i+1)) */
#define QCS_POS_YPOS         2
#define QCS_POS_ZPOS         3   /* z, nm (AS3D) */
#define QCS_POS_SIGX         4   /* sigma x, pixel */
#define QCS_POS_SIGY         5
#define QCS_POS_TPHO         6   /* total photon */
#define QCS_POS_BAKG         7   /* background photon per pixel */
#define QCS_POS_PSNR         8   /* peak SNR */
#define QCS_POS_CRBX         9   /* CRLB x, nm */
#define QCS_POS_CRBY         10  /* CRLB y, nm */
#define QCS_POS_FRME         11  /* frame number */

/* ---- ROI record: uint16[R*(R+1)] = R*R raw pixels row-major, then
 *      [Xc, Yc_in_frame, frame % 65536, frame / 65536, FitType, 0...]
 *      (DLL/SubregionExtraction/LDROIExtraction_ROIExtraction.cu:257-302) */
#define QCS_ROI_WORDS(R)     ((R) * ((R) + 1))
#define QCS_ROITYPE_SINGLE   0
#define QCS_ROITYPE_MULTI    1

/* ---- WLE pre-estimation record: float[6] per ROI
 *      (DLL/WLEParaEstimation/WLEParaEstimation_Parameters.h:20-34) */
#define QCS_WLE_FLOATS       6
#define QCS_WLE_NEARDIST     0
#define QCS_WLE_SIGMAL       1
#define QCS_WLE_SIGMAR       2
#define QCS_WLE_SIGMAU       3
#define QCS_WLE_SIGMAD       4
#define QCS_WLE_FLUOTYPE     5
#define QCS_MOLTYPE_MLEFIT   0
#define QCS_MOLTYPE_WLEFIT   1
#define QCS_MOLTYPE_MULTIFIT 2

/* ---- capacities (bfgs_CommonPara.h:315-316) */
#define QCS_POINT_NUM_TH     20480
#define QCS_MAX_POINT_NUM    (20480 * 5)

/* ---- rendering modes (DLL/ImgRender/ImageRender_base.h:36-39) */
#define QCS_RENDER_PEAKPHOTON_CRLB   0
#define QCS_RENDER_1PHOTON_CRLB      1
#define QCS_RENDER_1PHOTON_FIXED     2
#define QCS_RENDER_PEAKPHOTON_1PIXEL 3

/* ---- fit arithmetic.  EXACT evaluates the reference schedule with the library's own
 *      deterministic FP32 exp/log (bit-identical to the CPU oracle); FAST uses the gfx950
 *      hardware exp2/log2 instructions (same schedule, tolerance-level agreement). */
#define QCS_MATH_EXACT       0
#define QCS_MATH_FAST        1

/* Run configuration.  Mirrors LocalizationPara (DLL/bfgsMLE_Loc_LD/bfgs_base.h:26-110)
 * minus the GUI-only fields; defaults are LocalizationPara::LocalizationPara()
 * (DLL/bfgsMLE_Loc_LD/bfgs_base.cu:18-86). */
typedef struct qcs_params {
    /* camera */
    float Offset;            /* ADU */
    float KAdc;              /* e-/ADU */
    float QE;                /* e-/photon */
    float ReadNoise_e;       /* e- */
    /* localization */
    int   ROISize;           /* odd, 5..17 (5..13 with MultiEmitterFitEn, as the reference's kernels) */
    int   LocType;           /* QCS_LOCTYPE_* */
    int   MultiEmitterFitEn;
    int   WLEEn;
    int   ConsecFitEn;       /* session: merge molecules over consecutive frames (qcs_consec_*) */
    float ConsecFit_DistanceTh_nm;
    /* rendering */
    float PixelSize;         /* nm */
    float PixelZoom;
    float SNR_th;
    int   ColorMode_3D;
    /* image */
    int   ImageWidth;
    int   ImageHigh;
    int   TotalFrameNum;
    int   SRImageWidth;
    int   SRImageHigh;
    /* astigmatic 3D calibration: z = sum p_k d^k, d = sigx^2 - sigy^2 (px^2) */
    float MinZDepth;
    float MaxZDepth;
    float ZDepthCorrFactor;
    float p4_XGY, p3_XGY, p2_XGY, p1_XGY, p0_XGY;
    float p4_XLY, p3_XLY, p2_XLY, p1_XLY, p0_XLY;
    /* rendering selection (ImageRenderData_TypeDef::FluoRenderTop arguments) */
    int   RenderingMode;     /* QCS_RENDER_* */
    float FixedLocPrec;      /* nm, for QCS_RENDER_1PHOTON_FIXED */
    /* library controls (no reference counterpart) */
    int   FitMath;           /* QCS_MATH_EXACT (default) or QCS_MATH_FAST */
    int   MaxROIPerBatch;    /* ROI list capacity per extraction batch; default QCS_MAX_POINT_NUM */
    int   BatchedImageNum;   /* frames stacked per extraction batch (IJJ QC_STORM_.java:138-142); default 1 */
    int   FlushROINum;       /* session: ROIs accumulated before a fit; 0 = PointNumTh (20480; /4 for AS3D) */
} qcs_params;

/* Fill `p` with LocalizationPara() defaults (bfgs_base.cu:18-86) and the library controls. */
void qcs_params_default(qcs_params *p);
/* SRImageWidth/High = ((int)(W*PixelZoom)+3)/4*4 (bfgs_base.cu:89-98). */
void qcs_params_update_sr_size(qcs_params *p);

/* Thread-local text of the last error. */
const char *qcs_last_error(void);
/* Library version string. */
const char *qcs_version(void);

/* Test hook (no reference counterpart): the n-th device or pinned-host allocation the library
 * makes from now on fails as if the memory were exhausted (QCS_ERR_NOMEM); n < 0 disarms it.
 * Lets tests check that every entry point reports an allocation failure and leaves its context
 * usable. */
void qcs_test_fail_alloc_after(int n);

/* Number of HIP devices; <0 on runtime error.  Replaces CudaDeviceNum (REL/bfgs_MLE_dll.h:806-821). */
int qcs_device_count(void);

/* ========================================================================================
 * Stage API (device pointers).  These replace the methods of the static-library classes the
 * CPPDLL calls (REL/bfgs_MLE_dll.h:222-373), operating on HBM-resident data.  A context owns
 * the per-configuration scratch (filter taps, capacity buffers); it is not re-entrant: use one
 * context per host thread.
 * ====================================================================================== */
typedef struct qcs_ctx qcs_ctx;

/* Create a context bound to the current HIP device.  Replaces the Init(LocalizationPara&)
 * methods (LDROIExtraction.cu:219-286, bfgs_top.cu:234-293, ZeroLocalizationsRemove.cu:104-122). */
int qcs_ctx_create(const qcs_params *p, qcs_ctx **out);
/* Re-configure; reallocates only when sizes change (IJ/LocResource.cpp:219-224 re-init rule). */
int qcs_ctx_set_params(qcs_ctx *ctx, const qcs_params *p);
int qcs_ctx_destroy(qcs_ctx *ctx);

/* ROI extraction over `nframes` uint16 frames of ImageWidth x ImageHigh stacked vertically
 * (LDROIExtractData_TypeDef::ExtractMolecules, LDROIExtraction.cu:30-80, with WLE
 * pre-estimation WLEParaEstimation.cpp:20-40).  Frames are processed in batches of
 * BatchedImageNum; the candidate list of each batch is in raster order (frame, y, x).
 * Writes up to `cap` ROI records to d_rois and their WLE records to d_wle; *n_out receives
 * the number written.  Returns QCS_ERR_CAPACITY (with *n_out = cap) when more were found;
 * *n_found (may be NULL) receives the number found. */
int qcs_extract_dev(qcs_ctx *ctx, const uint16_t *d_frames, int nframes, int start_frame,
                    uint16_t *d_rois, float *d_wle, int cap, int *n_out, int *n_found,
                    void *stream);

/* WLE pre-estimation only (WLEParameterEstimation_TypeDef::WLEParameterEstimate,
 * WLEParaEstimation.cpp:20-40): PSF widths, per-frame nearest-neighbour distance, type. */
int qcs_wle_dev(qcs_ctx *ctx, const uint16_t *d_rois, int n, float *d_wle, void *stream);

/* BFGS MLE fitting + CRLB (LDLocData_TypeDef::BFGS_MLELocalization, bfgs_top.cu:26-204).
 * Row i of d_loc (12 floats) belongs to ROI i; in multi-emitter mode accepted second
 * emitters are appended at rows >= n.  `loc_cap` rows must be available (>= n; 2n covers
 * every append).  *n_rows receives the number of rows written (oValidFluoNum).
 * d_wle (n WLE records, qcs_wle_dev) is required: the fitter reads each ROI's type and PSF
 * width estimate from it. */
int qcs_fit_dev(qcs_ctx *ctx, const uint16_t *d_rois, const float *d_wle, int n,
                float *d_loc, int loc_cap, int *n_rows, void *stream);

/* Allocation-free qcs_fit_dev (no counterpart in the reference).  For callers that cannot
 * allocate device memory: code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc.  Same rows as qcs_fit_dev on the same inputs; the multi-emitter cascade's
 * scratch comes from a fixed file-scope device array (14 MiB per GPU, held from the first call on
 * that GPU until the process ends) instead of the context's buffers.  At most
 * QCS_NOALLOC_MAX_ROIS ROIs per call (QCS_ERR_CAPACITY beyond).  Nothing is waited for: the row
 * count goes to the device int d_n_rows (may be NULL), and qcs_fit_stats_get is not updated.
 * The context must belong to the current device.  Calls that overlap on one GPU share the array
 * and race: serialise them. */
#define QCS_NOALLOC_MAX_ROIS QCS_MAX_POINT_NUM
int qcs_fit_dev_noalloc(qcs_ctx *ctx, const uint16_t *d_rois, const float *d_wle, int n,
                        float *d_loc, int loc_cap, int *d_n_rows, void *stream);

/* Fit-stage counters of the last qcs_fit_dev call (FitRatio_* of bfgs_top.cu:68-173). */
typedef struct qcs_fit_stats {
    int   n_in, n_single, n_2e, n_3e, n_rejected, n_rows;
    float FitRatio_1E, FitRatio_2E, FitRatio_3E;
    float FitRatio_Final_1E, FitRatio_Final_2E, FitRatio_Final_3E, FitRatio_Final_4E;
} qcs_fit_stats;
int qcs_fit_stats_get(qcs_ctx *ctx, qcs_fit_stats *out);

/* Stage timing with HIP events recorded on the launch stream (the reference prints clock()
 * timers per stage, IJ/OnlineLocalizationLD.cpp:97-106,345-354).  Enabling resets the totals;
 * qcs_timing_get waits for the recorded events.  fit_1e covers the one-emitter kernel launches
 * (fit_1e_launches of them, fit_1e_rois ROIs), fit_me the multi-emitter ones. */
typedef struct qcs_timing {
    double extract_ms, fit_1e_ms, fit_me_ms, crlb_ms, compact_ms, render_ms;
    long long fit_1e_launches, fit_1e_rois, fit_me_launches;
    double cfit_ms, nn_ms;               /* converged fitting (qcs_cfit_dev, qcs_fit_mle_dev, warm path), network */
    long long cfit_launches, cfit_rois, nn_launches;
    double drift_cc_ms;                  /* drift correction: the cross-correlation kernel */
    long long drift_cc_launches, drift_cc_pixels;   /* launches; SR pixels correlated (47 x 47 shifts each) */
} qcs_timing;
int qcs_timing_enable(qcs_ctx *ctx, int on);
int qcs_timing_get(qcs_ctx *ctx, qcs_timing *out);

/* CRLB only (LocPrecCalc_GaussianCRLB.cu:23-40): fills CRBX/CRBY of rows [0, n). */
int qcs_crlb_dev(qcs_ctx *ctx, float *d_loc, int n, void *stream);

/* Frame-sorted removal of rejected rows (ZeroLocalizationsRemovel_TypeDef::RemoveZero_SortFrame,
 * ZeroLocalizationsRemove.cu:42-61,161-186): keeps rows with (int)PPho > 1 whose frame lies in
 * [first, end], ordered by frame, original order within a frame.  *n_out receives the count. */
int qcs_compact_sorted_dev(qcs_ctx *ctx, const float *d_loc, int n, int first_frame, int end_frame,
                           float *d_out, int *n_out, void *stream);
/* Unsorted removal (RemoveZeroWithoutSortFrame, ZeroLocalizationsRemove.cu:64-102): keeps rows
 * with PPho > 1.0f in original order. */
int qcs_compact_dev(qcs_ctx *ctx, const float *d_loc, int n, float *d_out, int *n_out, void *stream);

/* Splat-render rows into the float SR image d_img[SRImageHigh][SRImageWidth] (accumulating),
 * d_zimg (3D z-weighted plane, may be NULL for 2D) -- FluoRenderWithPrec_2D/3D
 * (ImageRender_2d.cu:55-155, ImageRender_3d.cu:60-147). */
int qcs_render_dev(qcs_ctx *ctx, const float *d_loc, int n, float *d_img, float *d_zimg, void *stream);
/* The same rendering into planes whose previous contents are replaced: the result equals clearing
 * d_img (and d_zimg, 3D) and calling qcs_render_dev (the same sums; float atomics add in the order
 * they happen to run, as in qcs_render_dev), in one write pass over the planes (no clear, no read).  No counterpart in the reference: for callers that draw a fresh image per
 * batch of rows, such as a per-batch display or benchmark step. */
int qcs_render_fresh_dev(qcs_ctx *ctx, const float *d_loc, int n, float *d_img, float *d_zimg, void *stream);

/* The colour-coded 3D image (ImgRender_GetSaveImg_3D / RenderSaveImg_3D,
 * ImgRender/ImageRender_3d.cu:238-403, 424-590).  d_img / d_zimg: the SRImageWidth x SRImageHigh
 * intensity and z planes of qcs_render_dev.  The brightness scale is the intensity below which 97%
 * of the pixels with intensity in (0, 12000) lie (2048-bin histogram) times bright_ratio
 * (*max_intensity, may be NULL); a pixel's value is min(intensity * 120 / scale, 100), its hue
 * (z plane / intensity) mapped over [MinZDepth, MaxZDepth] to 240..0 (ColorMode_3D 0, blue to
 * red) or 0..240 (1), saturation 100; zero-intensity pixels are black.  d_out: 3 (BGR, RGB) or 4
 * (BGRA, RGBA; alpha 0xff) bytes per pixel. */
#define QCS_RGB_BGR  0   /* RGBImage_EncodeMode_3B_BGR */
#define QCS_RGB_RGB  1
#define QCS_RGB_BGRA 2   /* the ImageJ plug-in's 3D display image */
#define QCS_RGB_RGBA 3
int qcs_render_rgb_dev(qcs_ctx *ctx, const float *d_img, const float *d_zimg, int mode, float bright_ratio, uint8_t *d_out,
                       float *max_intensity, void *stream);

/* ========================================================================================
 * Converged multi-emitter fitting (no counterpart in the reference, whose fitters run a fixed
 * iteration schedule), K = 1..3 emitters, Poisson MLE by BFGS with analytic gradient run to
 * convergence (step below 1e-4 in scaled units, at most 100 iterations), ROISize 7, 9 or 11:
 *   LocType 0 (2D):      mu = sum_k N_k/(2 pi s^2) exp(-d_k^2/(2 s^2)) + b, emitters sharing s,
 *                        theta = [N_k, x_k, y_k]_k, s, b (P = 3K + 2);
 *   LocType 1 (AS3D):    mu = sum_k N_k/(2 pi sx_k sy_k) exp(-(dx^2/(2 sx_k^2) + dy^2/(2 sy_k^2))) + b,
 *                        theta = [N_k, x_k, y_k, sx_k, sy_k]_k, b (P = 5K + 1, 16 at K = 3); rows
 *                        carry SigX/SigY = sx/sy and z from the p*_XGY / p*_XLY polynomial of
 *                        sx^2 - sy^2 times ZDepthCorrFactor, as the reference-schedule fitter does.
 * ====================================================================================== */
#define QCS_CFIT_INIT_HEUR  0   /* centroid / edge-background heuristic */
#define QCS_CFIT_INIT_GIVEN 1   /* d_init holds n x P starting vectors */
#define QCS_CFIT_INIT_ADD   2   /* d_init holds the (K-1)-emitter fits; add one at the largest residual */
#define QCS_CFIT_CONVERGED  1   /* info[1] flags */
#define QCS_CFIT_MAXITER    2
#define QCS_CFIT_STALLED    4
#define QCS_CFIT_FLAT       8   /* converged by ftol (set together with QCS_CFIT_CONVERGED) */
#define QCS_CFIT_PINNED    16   /* ended by pin_stop (see qcs_tuning) */
/* Fit model K to every ROI.  theta (n x P): N_k photons, x_k, y_k px (pixel edges), widths px,
 * b photons/px in the order above.  info (n x 4, may be NULL): iterations, flags, loss
 * evaluations, 0.  ll (n, may be NULL): FP64 log-likelihood sum(q ln mu - mu) at the fit. */
int qcs_cfit_dev(qcs_ctx *ctx, int K, const uint16_t *d_rois, int n, int init_mode, const float *d_init,
                 float *d_theta, int *d_info, double *d_ll, void *stream);
/* Tuning of the converged fits and of the warm start (per context). */
typedef struct qcs_tuning {
    int   max_iter[3];      /* BFGS iteration caps for K = 1, 2, 3 (default 100, 40, 40) */
    float tol;              /* convergence: quasi-Newton step below tol in every scaled coordinate (1e-4) */
    float warm_pmin;        /* warm start: minimum predicted-count probability, else plain path (0.8) */
    int   queue;            /* 1: persistent work-queue kernels (default); 0: one-pass kernels, a
                             * wave per 64 ROIs (same results) */
    int   waves_per_cu;     /* persistent waves per CU; 0 = as many as fit (default) */
    float ftol;             /* multi-emitter fits (K >= 2) also converge when an accepted step
                             * lowers the negative log-likelihood by less than ftol (1e-3; 0 = off) */
    int   active_set;       /* 1: parameters held at a box bound with an outward gradient are left
                             * out of the quasi-Newton step (default 0) */
    int   pin_stop;         /* 1: an elliptical three-emitter fit ends (QCS_CFIT_PINNED) after 10
                             * accepted iterations in a row with a photon count, position or width
                             * at its box bound, a point the selection rejects (default 0).  Measured
                             * on 2^18 9x9 astigmatic ROIs: the elliptical path 1.5x faster, count
                             * accuracy at three true emitters 0.33 instead of 0.36 */
} qcs_tuning;
/* ftol, active_set and max_iter apply as set to the 2D model.  Fits of the elliptical model
 * (LocType AS3D: theta = [N, x, y, sx, sy]_k, b) always run with the active set, without the
 * ftol stop and with at least 100 iterations. */
int qcs_tuning_get(qcs_ctx *ctx, qcs_tuning *out);
int qcs_tuning_set(qcs_ctx *ctx, const qcs_tuning *t);
/* The plain multi-emitter path: the emitter count follows the FP64 log-likelihood-ratio rule
 * (chi^2(3) at p = 1e-3) among valid models (every emitter >= 50 photons inside the ROI): K = 1
 * if its fit is valid, K = 2 if valid and (K = 0 or ll2 - ll1 > 8.14), K = 3 if K = 2 and valid
 * and ll3 - ll2 > 8.14.  K = 1 starts from the centroid heuristic, K = 2 from the K = 1 fit plus
 * an emitter at the largest residual, K = 3 likewise from K = 2 and only for ROIs at K = 2 (the
 * only ones the rule can move to 3).  3 rows per ROI (rows 3j..3j+2, unused rows zero) with the
 * FP64 full-Fisher CRLB in CRLBx/CRLBy.  d_sel (n, may be NULL) receives the chosen count. */
int qcs_fit_mle_dev(qcs_ctx *ctx, const uint16_t *d_rois, int n, int kmax, float *d_loc, int *d_sel, void *stream);
/* qcs_fit_mle_dev, then waits and fills h_stats (9 long longs) as qcs_fit_warm_dev does */
int qcs_fit_plain_stats_dev(qcs_ctx *ctx, const uint16_t *d_rois, int n, int kmax, float *d_loc, int *d_sel, long long *h_stats,
                            void *stream);
/* One set of same-size ROIs in a mixed-size call. */
typedef struct qcs_roi_set {
    int roi_size;             /* 7, 9 or 11 */
    int n;                    /* ROIs in the set; with d_n: the set's capacity */
    uint16_t *d_rois;         /* n ROI records of roi_size (written by qcs_extract_mixed_dev) */
    float *d_loc;             /* 3n rows, rows 3j..3j+2 belong to ROI j (as qcs_fit_mle_dev) */
    int *d_sel;               /* optional: chosen emitter count per ROI */
    int *d_n;                 /* optional: device int holding the set's ROI count (<= n); written by
                               * qcs_extract_mixed_dev, read by qcs_fit_mle_mixed_dev, so a set flows
                               * from extraction into fitting with no host round trip */
} qcs_roi_set;
/* qcs_fit_mle_dev over up to three sets of different ROI sizes: one persistent work-queue launch
 * per emitter count covers every set (largest ROIs first); the context's ROISize is not used.
 * Results equal per-set qcs_fit_mle_dev calls.  A set with d_n fits its first *d_n ROIs (grids
 * sized for n); nothing is waited for. */
int qcs_fit_mle_mixed_dev(qcs_ctx *ctx, const qcs_roi_set *sets, int nsets, int kmax, void *stream);

/* Mixed-size ROI extraction for the converged fitter (no counterpart in the reference, which
 * cuts one ROI size per run, LDROIExtraction_ROIExtraction.cu:257-302).  Detection,
 * classification and WLE as qcs_extract_dev at ROI size 7 (the context's other parameters),
 * at most `cap` candidates per call in raster order; each candidate then gets a size in
 * {7, 9, 11}:
 *   s = min(SigmaL, SigmaR, SigmaU, SigmaD) of its 7 x 7 WLE record; Rpsf = 7 (s <= 1.4),
 *   9 (s <= 1.9) or 11; with dmin the Chebyshev distance (1..5, same frame) of the nearest other
 *   detector mark, Rnb = 2 dmin + 3 for dmin 3 or 4, else 7; R = max(Rpsf, Rnb), reduced by 2
 *   while the R x R window holds more than 2 other marks, then while it leaves the frame.
 * sets[0..2] (roi_size 7, 9, 11 in that order, d_n required) receive the records of their
 * candidates in raster order, at most sets[k].n each, and the counts in *sets[k].d_n.
 * n_counts (3 ints, may be NULL) receives the counts after a stream synchronisation; with
 * NULL nothing is waited for. */
int qcs_extract_mixed_dev(qcs_ctx *ctx, const uint16_t *d_frames, int nframes, int start_frame, int cap,
                          const qcs_roi_set *sets, int *n_counts, void *stream);

/* Frames to localizations through the converged multi-emitter fitter (no counterpart in the
 * reference; its frame-driven multi-emitter flow is IJ/OnlineLocalizationLD.cpp:173-232 with the
 * fixed-schedule cascade): qcs_extract_mixed_dev (at most cap candidates, every set of capacity
 * cap), qcs_fit_mle_mixed_dev with emitter counts up to kmax, then the rows of every chosen
 * emitter in candidate order -- frame, row, column of the detection, then emitter -- so the
 * output is frame-sorted.  Rows carry the FP64 full-Fisher CRLB.  d_n_out (4 device ints)
 * receives the rows written (at most out_cap), the rows produced, the ROIs fitted and the
 * candidates found.  Nothing is waited for and
 * every count stays on the device, so the call can be captured in a HIP graph. */
int qcs_loc_mle_frames_dev(qcs_ctx *ctx, const uint16_t *d_frames, int nframes, int start_frame, int cap, int kmax,
                           float *d_out, int out_cap, int *d_n_out, void *stream);

/* ========================================================================================
 * Learned warm start (no counterpart in the reference): a 3-conv + 2-layer MLP on 7x7 ROIs
 * (62,192 parameters, BF16 weights on MFMA, FP32 accumulate, FP32 master weights and Adam),
 * trained on the GPU by this library's forward and backward kernels on simulator ROIs labelled
 * by the plain path (qcs_label_dev).  It predicts the emitter count (1-3) and the starting
 * vector of the converged fit.
 * ====================================================================================== */
typedef struct qcs_net qcs_net;
/* number of FP32 parameters (layout: W1[32][16] b1 W2[64][288] b2 W3[64][576] b3 W4[64][80] b4
 * W5[16][64] b5; conv1 taps 9..15 are zero padding) */
int qcs_net_param_count(void);
/* He-uniform initialisation from `seed` (deterministic); roi_size must be 7 */
int qcs_net_create(int roi_size, uint64_t seed, qcs_net **out);
int qcs_net_destroy(qcs_net *net);
int qcs_net_get_params(qcs_net *net, float *w);
/* load parameters (resets the optimiser state) */
int qcs_net_set_params(qcs_net *net, const float *w);
/* outputs [n][16]: 3 count logits, (x/7, y/7, N / (2 pi scale)) for 3 slots, s - 1.3,
 * (b - bhat) / scale; d_norm (n x 2, may be NULL): per-ROI (scale, bhat) of the input
 * normalisation.  One fused kernel (16 ROIs per wave, activations in registers). */
int qcs_net_forward_dev(qcs_net *net, const qcs_params *p, const uint16_t *d_rois, int n, float *d_out, float *d_norm,
                        void *stream);
/* the same outputs through the training path's layer-by-layer GEMMs (activations in HBM) */
int qcs_net_forward_layers_dev(qcs_net *net, const qcs_params *p, const uint16_t *d_rois, int n, float *d_out, float *d_norm,
                               void *stream);
/* labels [n][16]: K (0 = ignore), (x, y, N) x 3 (ROI pixels, photons, sorted by N), s, b.
 * Gradient of the mean loss (count cross-entropy + squared parameter errors) into h_grad
 * (qcs_net_param_count floats, may be NULL); *loss (may be NULL) receives the mean loss. */
int qcs_net_grad_dev(qcs_net *net, const qcs_params *p, const uint16_t *d_rois, const float *d_labels, int n, float *h_grad,
                     float *loss, void *stream);
/* one Adam step (beta 0.9 / 0.999) on the mean loss over the n ROIs */
int qcs_net_train_step_dev(qcs_net *net, const qcs_params *p, const uint16_t *d_rois, const float *d_labels, int n, float lr,
                           float *loss, void *stream);
/* labels from the plain path (qcs_fit_mle_dev's fits and model selection); d_sel (n, may be
 * NULL) receives the selected counts */
int qcs_label_dev(qcs_ctx *ctx, const uint16_t *d_rois, int n, float *d_labels, int *d_sel, void *stream);
/* The warm-started path (7x7 ROIs): the selection rule of qcs_fit_mle_dev with the network
 * choosing starts and pruning fits.  Every one-emitter fit starts from the network's brightest
 * emitter; a count the network predicts with probability >= warm_pmin caps the ROI's count
 * (predicted singles skip the multi-emitter fits, predicted doubles the three-emitter fit); the
 * others get the two-emitter fit from the network's two brightest emitters (where it predicts
 * two or more) or from the plain start, and K = 3 runs as on the plain path.  A warm
 * fit that does not converge or ends at a higher loss than its plain start (the heuristic, or
 * the K = 1 fit plus an emitter at the largest residual) is redone from that plain start.  Rows
 * as qcs_fit_mle_dev.  h_stats (9 long longs, may be NULL; the call then waits): ROIs per final
 * count [3], fits redone from the plain start, iterations, loss evaluations (all fits), ROIs
 * whose count the network capped, ROIs the network put at 2 and at 3 emitters. */
int qcs_fit_warm_dev(qcs_ctx *ctx, qcs_net *net, const uint16_t *d_rois, int n, float *d_loc, int *d_sel,
                     long long *h_stats, void *stream);

/* ========================================================================================
 * Statistics of a batch of localization rows (FluoStatisticData_TypeDef::GetStatisticalInf,
 * StatisticsInfo/statisticsInfo_top.cu:61-128; row kernel statisticsInfo.cu:19-91; activation
 * density statisticsInfo_top.cu:130-205, statisticsInfo_densityCalc.cu:17-107; helpers
 * statisticsInfo_rc.cu:17-152).  Rows: TotPho > 30, Bkg > 2, PSNR >= 3, PSF width
 * (SigX + SigY) / 2 >= 0.01 px and 0.01 < sqrt(CRLBx^2 + CRLBy^2) < 55 nm count as valid and
 * enter four histograms and the sums of the means.  The activation density comes from the
 * fractions of molecules with no and with exactly one neighbour within 10 px (20 px for 3D) in
 * the same frame, over the frames between the first and last row's frames (both ends excluded
 * when more than 3 frames), counting molecules 5 px or more inside the image.  The sums are FP64
 * (the reference sums with float atomics in arbitrary order); the call waits for its stream.
 * ====================================================================================== */
#define QCS_STAT_BINS 1600   /* StatInf_Hist_DatLenMax */
typedef struct qcs_loc_stats {
    int    hist_total_photon[QCS_STAT_BINS];   /* bins of 25 photons */
    int    hist_loc_prec[QCS_STAT_BINS];       /* sqrt(CRLBx^2 + CRLBy^2), bins of 0.15 nm */
    int    hist_peak_snr[QCS_STAT_BINS];       /* bins of 0.125 */
    int    hist_psf_width[QCS_STAT_BINS];      /* (SigX + SigY) / 2, bins of 0.0125 px */
    int    valid;                              /* rows that passed the filter */
    double sum_loc_prec, sum_peak_snr, sum_psf_width, sum_background;
    float  mean_total_photon;                  /* centre of mass of the photon histogram over bins >= 10% of its maximum */
    float  mean_loc_prec, mean_peak_snr, mean_background, mean_psf_width;
    float  mean_psf_width_ctl;                 /* mean_psf_width * -mean_peak_snr (lower is better) */
    int    density_first_frame, density_last_frame;
    int    density_total, density_n0, density_n1;   /* molecules considered; with 0 and with 1 neighbour */
    float  loc_density_2d;                     /* activation density, 1/um^2; 0 unless density_total > 50 */
} qcs_loc_stats;
int qcs_loc_stats_dev(qcs_ctx *ctx, const float *d_loc, int n, qcs_loc_stats *out, void *stream);

/* ========================================================================================
 * Consecutive-frame merging (ConsecutiveFit_TypeDef::ConsecutiveFit_WeightedAvg,
 * ConsecutiveFit/ConsecutiveFit_top.cu:21-101; kernels ConsecutiveFit.cu:18-220).  A qcs_consec
 * holds back the rows of each call until the next one.  A call links every row of [held rows,
 * d_rows] to the closest row of the next frame within ConsecFit_DistanceTh_nm / PixelSize pixels
 * (positions truncated to integers, squared distance strictly below the squared threshold, ties
 * to the lower row index; the reference's scan starts at the row's 32-row block, which matters
 * only for rows of frame 0), replaces each chain's first row by the 10/CRLBx-weighted average of
 * X, Y, Z, SigX, SigY with summed PPho, TotPho, Bkg and PSNR = sqrt(QE) PPho / sqrt(PPho + Bkg),
 * and zeroes the chain's other rows.  The CRLB (qcs_crlb_dev) of the held rows is recomputed.
 * d_out receives the held rows (is_end = 0; the call's rows are then held) or all rows
 * (is_end = 1; nothing is held), zeroed rows included: remove them with qcs_compact_dev, as the
 * reference does.  *n_out = number of rows written; out_cap must cover it.  d_rows must be sorted
 * by frame (rows of frame 0 may appear anywhere), as qcs_compact_sorted_dev writes them; other
 * orders fail with QCS_ERR_ARG and leave the held rows unchanged.  ontime_ratio (10 floats, may
 * be NULL): fraction of this call's chains (heads among d_rows) lasting 1, 2, ... >= 10 frames.
 * The parameters are the context's at call time; the context must outlive the qcs_consec.  The
 * call waits for `stream`.
 * ====================================================================================== */
#define QCS_ONTIME_BINS 10   /* MaxOnTimeConsecutiveNum */
typedef struct qcs_consec qcs_consec;
int qcs_consec_create(qcs_ctx *ctx, qcs_consec **out);
int qcs_consec_fit_dev(qcs_consec *cs, const float *d_rows, int n, int is_end, float *d_out, int out_cap, int *n_out,
                       float *ontime_ratio, void *stream);
int qcs_consec_held(const qcs_consec *cs);   /* rows held back (the reference's FluoNum_LastGroup) */
int qcs_consec_reset(qcs_consec *cs);        /* drop the held rows */
int qcs_consec_destroy(qcs_consec *cs);

/* ========================================================================================
 * Drift correction by cross-correlation of super-resolution images
 * (SRDriftCorrData_TypeDef::CorrectSampleShift, DriftCorrection_SRImageXCorr_top.cu:18-74; group
 * bookkeeping and interpolation DriftCorrection_SRImageXCorr_rc.cu:20-395; kernels
 * DriftCorrection_SRImageXCorr.cu:21-212).  d_rows: n frame-sorted rows (a result file's
 * contents), corrected in place.  The frame count is the last non-zero frame among the last 400
 * rows; the frames form groups of corr_frames overlapping by half (frames / corr_frames >= 1, at
 * most QCS_DRIFT_MAX_GROUPS groups).  Each group's rows with PSNR >= SNR_th and PPho > 1 are
 * rendered at 8x zoom of ImageWidth x ImageHigh as unit-weight Gaussians of sigma CRLBx (35 nm
 * outside (0.01, 35) nm); each group is correlated with group 0 over 47 x 47 SR-pixel shifts
 * centred on the previous group's integer shift, and the centre of mass of the peak region
 * (radius 12) gives its shift.  Shifts are interpolated linearly between group centres (from 0 at
 * frame 1) and subtracted from X and Y of rows of frames 1..frames (other rows are left as they
 * are).  shift_x / shift_y (may be NULL; shift_cap entries) receive the per-frame shifts in
 * pixels (index frame - 1), group_shift_xy (may be NULL; 2 x QCS_DRIFT_MAX_GROUPS floats) the
 * group shifts (x, y), *total_frames the frame count.  Uses 2 x 64 x ImageWidth x ImageHigh
 * floats of device scratch; the call waits for `stream`.
 * ====================================================================================== */
#define QCS_DRIFT_MAX_GROUPS 512   /* MaxCorrGroupNum */
int qcs_drift_correct_dev(qcs_ctx *ctx, float *d_rows, long long n, int corr_frames, float *shift_x, float *shift_y,
                          int shift_cap, int *total_frames, float *group_shift_xy, void *stream);

/* ========================================================================================
 * Per-batch HIP graphs (no counterpart in the reference).  One call of qcs_fit_mle_dev (net
 * NULL) or qcs_fit_warm_dev (statistics off) over fixed device buffers is captured into a HIP
 * graph; each qcs_graph_launch refits whatever d_rois holds at that time into d_loc / d_sel.  The
 * graph owns a private context with the caller's parameters and tuning (copied at creation), so
 * the scratch addresses it uses stay valid; create runs the path once eagerly (after waiting for
 * `stream`) to size that scratch.  Replays of one graph must not overlap.
 * ====================================================================================== */
typedef struct qcs_graph qcs_graph;
int qcs_graph_create_mle(qcs_ctx *ctx, qcs_net *net, const uint16_t *d_rois, int n, int kmax, float *d_loc, int *d_sel,
                         void *stream, qcs_graph **out);
/* One frame batch through qcs_loc_mle_frames_dev (mixed-size extraction, converged fits,
 * frame-sorted rows) captured over fixed buffers: each launch localizes the nframes frames then
 * in d_frames, numbered from the device int *d_start_frame as it is at run time, into d_out /
 * d_n_out.  A per-frame-batch pipeline writes each batch and its first frame number and launches
 * the graph; replays equal eager calls. */
int qcs_graph_create_frames(qcs_ctx *ctx, const uint16_t *d_frames, int nframes, const int *d_start_frame, int cap, int kmax,
                            float *d_out, int out_cap, int *d_n_out, void *stream, qcs_graph **out);
int qcs_graph_launch(qcs_graph *g, void *stream);
int qcs_graph_destroy(qcs_graph *g);

/* ========================================================================================
 * Autotuner of the multi-emitter paths (no counterpart in the reference).  On n sample ROIs it
 * runs the plain path fully converged (caps 100, ftol 0) as the reference, then times candidate
 * settings -- iteration caps, ftol, persistent waves per CU, persistent vs one-pass kernels, and
 * with a network warm_pmin -- with HIP events (trials runs each, after one warm-up run) and
 * scores each against the reference: the fraction of ROIs with the same emitter count and the
 * 99th percentile of max(|dx|, |dy|) (pixels) over the emitters of those ROIs.  The fastest plain
 * candidate with select_agree >= min_select_agree and dx_p99 <= max_dx_p99 sets the fitting
 * knobs, the fastest such warm candidate warm_pmin; the result becomes the context's tuning
 * (qcs_tuning_get).  d_loc (3n rows) and d_sel (n) are scratch for the runs.  table (may be NULL
 * with table_cap 0) receives up to table_cap candidates, *n_entries their number.
 * ====================================================================================== */
typedef struct qcs_autotune_opts {
    int   trials;             /* timed runs per candidate, 1..100 */
    float min_select_agree;   /* e.g. 0.99 */
    float max_dx_p99;         /* pixels, e.g. 0.02 */
} qcs_autotune_opts;
typedef struct qcs_autotune_entry {
    qcs_tuning tuning;
    float ms;                 /* mean time per call of n ROIs */
    float select_agree, dx_p99;
    int   warm;               /* 1: warm path (network) */
    int   within;             /* met both tolerances */
    int   chosen;
} qcs_autotune_entry;
int qcs_autotune_mle_dev(qcs_ctx *ctx, qcs_net *net, const uint16_t *d_rois, int n, float *d_loc, int *d_sel,
                         const qcs_autotune_opts *opts, qcs_autotune_entry *table, int table_cap, int *n_entries, void *stream);
/* building block: C[M][N] (fp32) = A[M][K] B[N][K]^T, bf16 inputs, K a multiple of 16 */
int qcs_dev_gemm_nt_bf16(const uint16_t *d_a, const uint16_t *d_b, float *d_c, int M, int N, int K, void *stream);

/* ========================================================================================
 * Host-pointer stage API.  Same semantics as the *_dev calls, with the H2D / D2H copies the
 * reference performs inside the class methods (e.g. bfgs_top.cu:42-47,201-202).
 * ====================================================================================== */
int qcs_fit_rois(qcs_ctx *ctx, const uint16_t *rois, const float *wle, int n,
                 float *loc, int loc_cap, int *n_rows);
int qcs_extract(qcs_ctx *ctx, const uint16_t *frames, int nframes, int start_frame,
                uint16_t *rois, float *wle, int cap, int *n_out, int *n_found);

/* ========================================================================================
 * Online API.  Mirrors the lm_* JNI entry points (IJ/QC_STORM_CPPDLL.h:15-174,
 * MM/QC_STORM_CPPDLL.cpp:38-924): frames are queued and localized by an internal worker
 * thread; localizations are frame-sorted and accumulated; the SR image is rendered on device.
 * ====================================================================================== */
typedef struct qcs_session qcs_session;

/* lm_SetLocPara / lm_SetLocPara3D / lm_SetImagePara + lm_StartLocThread. */
int qcs_session_create(const qcs_params *p, qcs_session **out);
/* lm_FeedImageData(short[], n): copies `nframes` frames into a bounded queue (blocks while the
 * queue holds `max_queued_frames` frames, the reference's back-pressure). */
int qcs_session_feed(qcs_session *s, const uint16_t *frames, int nframes);
/* lm_StopLocThread: process everything queued, then stop the worker. */
int qcs_session_finish(qcs_session *s);
/* lm_IsLocFinish: 1 when the worker has drained the queue after finish, else 0. */
int qcs_session_is_finished(qcs_session *s);
/* Localizations accumulated so far (frame-sorted).  Copies min(count, cap) rows into `out12`
 * (may be NULL to query) and returns the total count (<0 on error). */
long long qcs_session_fetch_locs(qcs_session *s, float *out12, long long cap);
/* lm_GetSMLMImage: copies the SRImageWidth*SRImageHigh float image. */
int qcs_session_get_sr_image(qcs_session *s, float *img);
/* lm_GetSMLMImage3D (IJ/QC_STORM_CPPDLL.cpp:283-309; the image of OnlineRendDisp.cpp:35-43): the
 * colour-coded 3D image of the accumulated rows, 4 bytes BGRA per SR pixel (qcs_render_rgb_dev,
 * bright ratio 1).  AS3D sessions only. */
int qcs_session_get_sr_image3d(qcs_session *s, uint8_t *bgra);
/* lm_GetMaxDispVal: max of (int) SR image value. */
int qcs_session_get_max_disp_val(qcs_session *s, int *val);
/* lm_ReleaseResource. */
int qcs_session_destroy(qcs_session *s);
/* The reference's LocFile (IJ/OnlineLocalizationLD.cpp:70-71, 283): every flush's final rows --
 * frame-sorted, consecutive-frame merged when ConsecFitEn -- are appended to `path` as raw
 * little-endian 12-float records (48 bytes each), by a writer thread of the session.  Call before
 * the first qcs_session_feed; the file is complete when qcs_session_is_finished returns 1 (a
 * write failure is reported as QCS_ERR_IO by the session calls). */
int qcs_session_set_output(qcs_session *s, const char *path);
/* FluoStatisticData_TypeDef after the last flush (GetStatisticalInf on the flush's rows,
 * statisticsInfo_top.cu:61-128) and, when ConsecFitEn, the flush's on-time ratios
 * (UpdateOntimeRatio).  Either pointer may be NULL. */
int qcs_session_get_stats(qcs_session *s, qcs_loc_stats *out, float *ontime_ratio);
/* SetLocDataFileName (IJ/LocResource.cpp:160-179): image name with the trailing '.', 't', 'i',
 * 'f' characters removed (CString::TrimRight(".tif")), then "_result<LocType+2>D<ROISize>_<M|S>"
 * and "_Consec<ConsecFit_DistanceTh_nm>nm" when ConsecFitEn, then ".txt".  Copies at most cap-1
 * characters and a terminating 0 into out; returns the full length (<0 on error). */
int qcs_result_file_name(const qcs_params *p, const char *image_name, char *out, int cap);

/* lm_LocBatchedImg (MM/QC_STORM_CPPDLL.cpp:857-924): synchronous extract + fit + frame-sorted
 * compaction of `nframes` frames; returns the number of rows (<0 on error) and copies up to
 * `cap` of them into out12.  Frames are numbered from `start_frame`. */
long long qcs_loc_batch_sync(qcs_ctx *ctx, const uint16_t *frames, int nframes, int start_frame,
                             float *out12, long long cap);
/* lm_LocBatchedImg with its return array (MM/QC_STORM_CPPDLL.cpp:857-924; the flow of
 * LocResource_ZDriftMeasure.h:42-70, statistics reset per call): as qcs_loc_batch_sync, then the
 * statistics of the batch's rows (qcs_loc_stats_dev) into `stats` (may be NULL) and the 8-float
 * summary into `info8` (may be NULL): [0] mean peak SNR, [1] -mean PSF width, [2] [0] + [1],
 * [3] [0] * [1], [4] rows, [5] activation density (0 when 50 or fewer molecules), [6], [7] 0. */
long long qcs_loc_batch_sync_info(qcs_ctx *ctx, const uint16_t *frames, int nframes, int start_frame,
                                  float *out12, long long cap, float *info8, qcs_loc_stats *stats);

/* ========================================================================================
 * Synthetic data (deterministic simulator; no reference counterpart).  Philox4x32-10 keyed
 * by (seed, unit, pixel); Poisson by inversion (lambda < 10) or PTRS (Hoermann 1993).
 * ====================================================================================== */
typedef struct qcs_sim_params {
    int      width, height;        /* ROI size (R x R) or frame size */
    float    offset, kadc, qe;     /* ADU = min(65535, floor(e/kadc + 0.5) + offset) */
    float    photons_min, photons_max;   /* log-uniform */
    float    sigma_min, sigma_max;       /* uniform (2D) */
    float    bg_min, bg_max;             /* uniform, photons/px */
    float    density_per_px;             /* frames: emitters per pixel per frame */
    float    center_jitter;              /* ROIs: centre uniform within +-jitter of R/2 */
    int      astig;                      /* 1: sigma_x/sigma_y from z (astigmatic model) */
    float    z_min, z_max, s0, c_nm, d_nm;   /* astigmatic model */
    int      n_per_unit_max;             /* ROIs: 1..n emitters (multi-emitter ROIs); frames: emitters
                                          * come in sites of 1..n (clusters) */
    float    min_sep, max_sep;           /* ROIs / sites with >1 emitter: separation range, px */
    float    density_max_per_px;         /* frames: when above density_per_px, each frame's density
                                          * is uniform in [density_per_px, density_max_per_px] */
} qcs_sim_params;
void qcs_sim_params_default(qcs_sim_params *sp);

/* Ground truth per emitter: [unit, x, y, z, photons, bg, sigx, sigy] (8 floats). */
#define QCS_TRUTH_FLOATS 8

/* Simulate `n` independent R x R ROI records (headers: Xc = Yc = 100, frame = first_frame + i)
 * and their ground truth (n_per_unit_max floats-of-8 per ROI, unused slots zero).  Host and
 * device versions produce identical records. */
int qcs_sim_rois_host(const qcs_sim_params *sp, uint64_t seed, int n, int first_frame,
                      uint16_t *rois, float *truth);
int qcs_sim_rois_dev(const qcs_sim_params *sp, uint64_t seed, int n, int first_frame,
                     uint16_t *d_rois, float *d_truth, void *stream);
/* Number of emitters the simulator places in frame `frame` (deterministic); truth (cap rows of
 * QCS_TRUTH_FLOATS, may be NULL) receives them. */
int qcs_sim_frame_emitters(const qcs_sim_params *sp, uint64_t seed, int frame, float *truth, int cap);
/* Simulate `nframes` frames width x height starting at frame index `first_frame`. */
int qcs_sim_frames_host(const qcs_sim_params *sp, uint64_t seed, int first_frame, int nframes,
                        uint16_t *frames);
int qcs_sim_frames_dev(const qcs_sim_params *sp, uint64_t seed, int first_frame, int nframes,
                       uint16_t *d_frames, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* QCSTORM_H */
