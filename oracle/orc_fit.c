/*
 * orc_fit.c -- FP32 and FP64 instantiations of the oracle fitters (TEST INFRASTRUCTURE).
 * The algorithm text is in orc_fit_impl.h; see there for the reference citations.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "orc.h"

float orc_exp32(float x);
float orc_log32(float x);
float orc_log32_u(float x, float c);
float orc_log32_ub(float x, float c);
int   orc_model_separable(void);

/* loss evaluations of the calling thread (the schedule's cost; read by the tests) */
static _Thread_local long g_loss_evals;
static void orc_count_loss_eval(void) { g_loss_evals++; }
long orc_loss_evals(int reset)
{
    long v = g_loss_evals;
    if (reset) g_loss_evals = 0;
    return v;
}

/* ---------------- FP32 ---------------- */
#define REAL        float
#define SUF(n)      n##_f32
#define FMA(a,b,c)  fmaf((a), (b), (c))
#define EXPR(x)     orc_exp32(x)
#define LOGR(x)     orc_log32(x)
#define LOGU(x,c)   orc_log32_u((x), (c))
#define LOGUB(x,c)  orc_log32_ub((x), (c))
#define SQRTR(x)    sqrtf(x)
#define FABSR(x)    fabsf(x)
#define SEPARABLE() orc_model_separable()
#define C(x)        ((float)(x))
#include "orc_fit_impl.h"
#undef REAL
#undef SUF
#undef FMA
#undef EXPR
#undef LOGR
#undef LOGU
#undef LOGUB
#undef SQRTR
#undef FABSR
#undef SEPARABLE
#undef C
#undef MAXP
#undef MAXR
#undef MAXPIX
#define M_GS2D_F 0
#define M_AS3D_F 1
#undef M_GS2D
#undef M_AS3D
#undef M_2E
#undef M_3E
#undef M_AS3D2E

/* ---------------- FP64 (same schedule, double precision, libm exp/log) ---------------- */
#define REAL        double
#define SUF(n)      n##_f64
#define FMA(a,b,c)  fma((a), (b), (c))
#define EXPR(x)     exp(x)
#define LOGR(x)     log(x)
#define LOGU(x,c)   (log(x) + (c))
#define LOGUB(x,c)  (log(x) + (c))
#define SQRTR(x)    sqrt(x)
#define FABSR(x)    fabs(x)
#define SEPARABLE() 0
#define C(x)        ((double)(x))
#include "orc_fit_impl.h"

static float *wle_or_compute(const qcs_params *p, const uint16_t *rois, const float *wle, int n, float **owned)
{
    *owned = NULL;
    if (wle) return (float *)wle;
    float *w = (float *)calloc((size_t)n * QCS_WLE_FLOATS + 1, sizeof(float));
    orc_wle(rois, n, p->ROISize, p->LocType, w);
    *owned = w;
    return w;
}

int orc_fit(const qcs_params *p, const uint16_t *rois, const float *wle, int n, float *loc,
            int loc_cap, orc_fit_stats *st)
{
    float *owned;
    const float *w = wle_or_compute(p, rois, wle, n, &owned);
    int r = fit_all_f32(p, rois, w, n, loc, loc_cap, st);
    free(owned);
    return r;
}

/* Single-emitter mode over ROIs in parallel (OpenMP, `threads` threads): the same per-ROI code as
 * orc_fit, used as the CPU baseline; returns rows written (= n) or -1 in multi-emitter mode. */
int orc_fit_omp(const qcs_params *p, const uint16_t *rois, const float *wle, int n, float *loc, int threads)
{
    if (p->MultiEmitterFitEn) return -1;
    const int R = p->ROISize, W = QCS_ROI_WORDS(R);
    const int model = (p->LocType == QCS_LOCTYPE_AS3D) ? M_AS3D_F : M_GS2D_F;
#pragma omp parallel for schedule(dynamic, 256) num_threads(threads)
    for (int i = 0; i < n; i++) {
        float snr;
        int valid;
        fit_1e_f32(p, model, rois + (long)i * W, wle + (long)i * QCS_WLE_FLOATS, loc + (long)i * QCS_REC_FLOATS, &snr, &valid);
    }
#pragma omp parallel for schedule(static) num_threads(threads)
    for (int i = 0; i < n; i++) crlb_rows_f32(p, loc + (long)i * QCS_REC_FLOATS, 1);
    return n;
}

int orc_fit_f64(const qcs_params *p, const uint16_t *rois, const float *wle, int n, double *loc,
                int loc_cap, orc_fit_stats *st)
{
    float *owned;
    const float *w = wle_or_compute(p, rois, wle, n, &owned);
    int r = fit_all_f64(p, rois, w, n, loc, loc_cap, st);
    free(owned);
    return r;
}

int orc_fit_params(const qcs_params *p, int model, const uint16_t *rois, const float *wle, int n,
                   float *theta, int iters)
{
    float *owned;
    const float *w = wle_or_compute(p, rois, wle, n, &owned);
    int r = fit_params_f32(p, model, rois, w, n, theta, iters);
    free(owned);
    return r;
}

int orc_fit_params_f64(const qcs_params *p, int model, const uint16_t *rois, const float *wle,
                       int n, double *theta, int iters)
{
    float *owned;
    const float *w = wle_or_compute(p, rois, wle, n, &owned);
    int r = fit_params_f64(p, model, rois, w, n, theta, iters);
    free(owned);
    return r;
}

float orc_loss(const qcs_params *p, int model, const uint16_t *roi, const float *wle, const float *theta)
{
    mdesc_f32 m;
    roictx_f32 rc;
    float x[11], t[11];
    get_desc_f32(model, &m);
    setup_f32(p, model, roi, wle, &rc, x, NULL);
    for (int i = 0; i < m.P; i++) t[i] = theta[i];
    return loss_f32(&m, &rc, t);
}

void orc_crlb(const qcs_params *p, float *loc, int n) { crlb_rows_f32(p, loc, n); }

/* Acceptance rules on a given fit vector (scaled units, as the fitters hold it). */
int orc_select_1e(const qcs_params *p, int model, const float *x, const uint16_t *roi, float *row)
{
    float snr;
    int valid;
    return select_1e_f32(p, model, x, roi, row, &snr, &valid);
}

int orc_select_me(const qcs_params *p, int model, const float *x, float nn, const uint16_t *roi, float *row, float *row2)
{
    return select_me_f32(p, model, x, nn, roi, row, row2);
}
void orc_crlb_f64(const qcs_params *p, double *loc, int n) { crlb_rows_f64(p, loc, n); }
