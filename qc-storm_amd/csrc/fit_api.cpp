// fit_api.cpp -- fitting stage entry points (LDLocData_TypeDef::BFGS_MLELocalization,
// DLL/bfgsMLE_Loc_top/bfgs_top.cu:26-204).
#include <string.h>
#include "qcs_internal.h"

using namespace qcs;

namespace qcs {
int fit_cascade(qcs_ctx *c, const uint16_t *d_rois, const float *d_wle, int n, float *d_loc, int loc_cap, int *n_rows,
                hipStream_t s);
}

extern "C" {

int qcs_fit_dev(qcs_ctx *c, const uint16_t *d_rois, const float *d_wle, int n, float *d_loc, int loc_cap, int *n_rows,
                void *stream)
{
    if (!c || n < 0 || (n > 0 && (!d_rois || !d_loc))) return fail(QCS_ERR_ARG, "invalid argument");
    if (n_rows) *n_rows = 0;
    memset(&c->stats, 0, sizeof(c->stats));
    c->stats.n_in = n;
    if (n == 0) return QCS_OK;
    if (loc_cap < n) return fail(QCS_ERR_CAPACITY, "loc_cap (%d) < number of ROIs (%d)", loc_cap, n);
    if (!d_wle) return fail(QCS_ERR_ARG, "WLE records are required (qcs_wle_dev computes them)");
    // the reference instantiates its 2- and 3-emitter kernels for ROI sizes 5..13 only
    // (bfgsMLE_Multi_Emitter_Fit.cu:44-75: "shared memory is not enough" for 15 and 17)
    if (c->p.MultiEmitterFitEn && c->p.ROISize > 13)
        return fail(QCS_ERR_ARG, "multi-emitter fitting supports ROISize 5..13 (got %d)", c->p.ROISize);
    hipStream_t s = (hipStream_t)stream;
    if (!c->p.MultiEmitterFitEn) {
        hipEvent_t ev;
        int rc = timer_begin(c->timer, s, &ev);
        if (rc) return rc;
        rc = fit_launch_1e(c->p, d_rois, d_wle, nullptr, n, nullptr, d_loc, nullptr, s);
        if (rc) return rc;
        if ((rc = timer_end(c->timer, TM_FIT1E, s, ev, n)) || (rc = timer_begin(c->timer, s, &ev))) return rc;
        rc = crlb_launch(c->p, d_loc, n, nullptr, s);
        if (rc) return rc;
        if ((rc = timer_end(c->timer, TM_CRLB, s, ev, n))) return rc;
        c->stats.n_single = n;
        c->stats.n_rows = n;
        c->stats.FitRatio_1E = 1.0f;
        c->stats.FitRatio_Final_1E = 1.0f;
        if (n_rows) *n_rows = n;
        return QCS_OK;
    }
    return fit_cascade(c, d_rois, d_wle, n, d_loc, loc_cap, n_rows, s);
}

int qcs_fit_rois(qcs_ctx *c, const uint16_t *rois, const float *wle, int n, float *loc, int loc_cap, int *n_rows)
{
    if (!c || n < 0 || (n > 0 && (!rois || !loc || !wle))) return fail(QCS_ERR_ARG, "invalid argument");
    if (n_rows) *n_rows = 0;
    if (n == 0) return QCS_OK;
    const int R = c->p.ROISize;
    const size_t rb = (size_t)n * QCS_ROI_WORDS(R) * sizeof(uint16_t);
    const size_t wb = (size_t)n * QCS_WLE_FLOATS * sizeof(float);
    const size_t lb = (size_t)loc_cap * QCS_REC_FLOATS * sizeof(float);
    int rc;
    if ((rc = c->h_rois_dev.ensure(rb)) || (rc = c->h_wle_dev.ensure(wb)) || (rc = c->h_loc_dev.ensure(lb))) return rc;
    QCS_HIP(hipMemcpy(c->h_rois_dev.ptr, rois, rb, hipMemcpyHostToDevice));
    QCS_HIP(hipMemcpy(c->h_wle_dev.ptr, wle, wb, hipMemcpyHostToDevice));
    int rows = 0;
    rc = qcs_fit_dev(c, c->h_rois_dev.as<uint16_t>(), c->h_wle_dev.as<float>(), n, c->h_loc_dev.as<float>(), loc_cap,
                     &rows, nullptr);
    if (rc) return rc;
    QCS_HIP(hipDeviceSynchronize());
    QCS_HIP(hipMemcpy(loc, c->h_loc_dev.ptr, (size_t)rows * QCS_REC_FLOATS * sizeof(float), hipMemcpyDeviceToHost));
    if (n_rows) *n_rows = rows;
    return QCS_OK;
}

}  // extern "C"

