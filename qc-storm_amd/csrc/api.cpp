// api.cpp -- C-ABI of libqcstorm: parameters, errors, contexts and the stage entry points.
#include <stdarg.h>
#include <stdio.h>
#include <string.h>
#include <atomic>
#include <new>
#include <string>
#include "qcs_internal.h"

namespace qcs {

static thread_local std::string g_err;

int fail(int code, const char *fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

int hip_fail(hipError_t e, const char *what, const char *file, int line)
{
    return fail(QCS_ERR_HIP, "%s failed: %s (%s:%d)", what, hipGetErrorString(e), file, line);
}

// test hook (qcs_test_fail_alloc_after): allocations left before one fails, -1 = never
static std::atomic<int> g_fail_alloc_after{-1};

static bool injected_alloc_failure()
{
    int v = g_fail_alloc_after.load();
    while (v >= 0) {
        if (g_fail_alloc_after.compare_exchange_weak(v, v - 1)) return v == 0;
    }
    return false;
}

int DevBuf::ensure(size_t b)
{
    if (b <= bytes && ptr) return QCS_OK;
    release();
    if (b == 0) return QCS_OK;
    if (injected_alloc_failure()) return fail(QCS_ERR_NOMEM, "hipMalloc(%zu) failed: injected by qcs_test_fail_alloc_after", b);
    hipError_t e = hipMalloc(&ptr, b);
    if (e != hipSuccess) {
        ptr = nullptr;
        bytes = 0;
        return fail(QCS_ERR_NOMEM, "hipMalloc(%zu) failed: %s", b, hipGetErrorString(e));
    }
    bytes = b;
    return QCS_OK;
}

void DevBuf::release()
{
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
}

int HostBuf::ensure(size_t b)
{
    if (b <= bytes && ptr) return QCS_OK;
    release();
    if (b == 0) return QCS_OK;
    if (injected_alloc_failure()) return fail(QCS_ERR_NOMEM, "hipHostMalloc(%zu) failed: injected by qcs_test_fail_alloc_after", b);
    hipError_t e = hipHostMalloc(&ptr, b, hipHostMallocDefault);
    if (e != hipSuccess) {
        ptr = nullptr;
        bytes = 0;
        return fail(QCS_ERR_NOMEM, "hipHostMalloc(%zu) failed: %s", b, hipGetErrorString(e));
    }
    bytes = b;
    return QCS_OK;
}

void HostBuf::release()
{
    if (ptr) (void)hipHostFree(ptr);
    ptr = nullptr;
    bytes = 0;
}

int validate_params(const qcs_params &p)
{
    if (p.ROISize < 5 || p.ROISize > 17 || (p.ROISize & 1) == 0)
        return fail(QCS_ERR_ARG, "ROISize must be odd in [5, 17] (got %d)", p.ROISize);
    if (p.LocType < QCS_LOCTYPE_GS2D || p.LocType > QCS_LOCTYPE_DH3D)
        return fail(QCS_ERR_ARG, "LocType %d invalid", p.LocType);
    if (!(p.KAdc > 0.0f) || !(p.QE > 0.0f)) return fail(QCS_ERR_ARG, "KAdc and QE must be positive");
    if (p.FitMath != QCS_MATH_EXACT && p.FitMath != QCS_MATH_FAST) return fail(QCS_ERR_ARG, "FitMath %d invalid", p.FitMath);
    if (p.ImageWidth < 0 || p.ImageHigh < 0) return fail(QCS_ERR_ARG, "negative image size");
    if (p.FlushROINum < 0) return fail(QCS_ERR_ARG, "FlushROINum must be >= 0 (got %d)", p.FlushROINum);
    return QCS_OK;
}

int timer_begin(Timer &t, hipStream_t s, hipEvent_t *ev)
{
    *ev = nullptr;
    if (!t.on) return QCS_OK;
    QCS_HIP(hipEventCreate(ev));
    QCS_HIP(hipEventRecord(*ev, s));
    return QCS_OK;
}

int timer_end(Timer &t, int cat, hipStream_t s, hipEvent_t ev, long long items)
{
    if (!t.on || !ev) return QCS_OK;
    hipEvent_t b;
    QCS_HIP(hipEventCreate(&b));
    QCS_HIP(hipEventRecord(b, s));
    if (t.npend == t.cap) {
        int nc = t.cap ? 2 * t.cap : 64;
        Timer::Pending *np = new (std::nothrow) Timer::Pending[nc];
        if (!np) return fail(QCS_ERR_NOMEM, "timer allocation failed");
        for (int i = 0; i < t.npend; i++) np[i] = t.pend[i];
        delete[] t.pend;
        t.pend = np;
        t.cap = nc;
    }
    t.pend[t.npend++] = {cat, ev, b, items};
    return QCS_OK;
}

static int timer_resolve(Timer &t)
{
    for (int i = 0; i < t.npend; i++) {
        Timer::Pending &p = t.pend[i];
        QCS_HIP(hipEventSynchronize(p.b));
        float ms = 0.0f;
        QCS_HIP(hipEventElapsedTime(&ms, p.a, p.b));
        t.ms[p.cat] += ms;
        t.launches[p.cat] += 1;
        t.items[p.cat] += p.items;
        (void)hipEventDestroy(p.a);
        (void)hipEventDestroy(p.b);
    }
    t.npend = 0;
    return QCS_OK;
}

}  // namespace qcs

using namespace qcs;

extern "C" {

void qcs_params_default(qcs_params *p)
{
    if (!p) return;
    memset(p, 0, sizeof(*p));
    // LocalizationPara::LocalizationPara (bfgs_base.cu:18-86)
    p->KAdc = 0.45f;
    p->Offset = 100.0f;
    p->QE = 0.72f;
    p->ReadNoise_e = 1.3f;
    p->ROISize = 7;
    p->LocType = QCS_LOCTYPE_GS2D;
    p->MultiEmitterFitEn = 0;
    p->WLEEn = 1;
    p->ConsecFit_DistanceTh_nm = 80.0f;
    p->ConsecFitEn = 0;
    p->PixelSize = 100.0f;
    p->PixelZoom = 5.0f;
    p->SNR_th = 5.0f;
    p->ColorMode_3D = 0;
    p->ImageWidth = 2048;
    p->ImageHigh = 2048;
    p->TotalFrameNum = 1;
    p->MinZDepth = -500.0f;
    p->MaxZDepth = 500.0f;
    p->ZDepthCorrFactor = 1.0f;
    p->p4_XGY = 0.0f; p->p3_XGY = 0.0f; p->p2_XGY = 0.0f; p->p1_XGY = 1.0f; p->p0_XGY = 0.0f;
    p->p4_XLY = 0.0f; p->p3_XLY = 0.0f; p->p2_XLY = 0.0f; p->p1_XLY = 1.0f; p->p0_XLY = 0.0f;
    p->RenderingMode = QCS_RENDER_PEAKPHOTON_CRLB;
    p->FixedLocPrec = 20.0f;
    p->FitMath = QCS_MATH_EXACT;
    p->MaxROIPerBatch = QCS_MAX_POINT_NUM;
    p->BatchedImageNum = 1;
    p->FlushROINum = 0;
    qcs_params_update_sr_size(p);
}

void qcs_params_update_sr_size(qcs_params *p)
{
    if (!p) return;
    // LocalizationPara::UpdateSRImageSize (bfgs_base.cu:89-98)
    int w = (int)((float)p->ImageWidth * p->PixelZoom);
    int h = (int)((float)p->ImageHigh * p->PixelZoom);
    p->SRImageWidth = (w + 3) / 4 * 4;
    p->SRImageHigh = (h + 3) / 4 * 4;
}

const char *qcs_last_error(void) { return g_err.c_str(); }

const char *qcs_version(void) { return "qcstorm-mi355x 0.1 (gfx950)"; }

int qcs_device_count(void)
{
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) {
        fail(QCS_ERR_HIP, "hipGetDeviceCount: %s", hipGetErrorString(e));
        return -1;
    }
    return n;
}

int qcs_ctx_create(const qcs_params *p, qcs_ctx **out)
{
    if (!p || !out) return fail(QCS_ERR_ARG, "null argument");
    *out = nullptr;
    int rc = validate_params(*p);
    if (rc) return rc;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return fail(QCS_ERR_NODEVICE, "no HIP device available");
    qcs_ctx *c = new (std::nothrow) qcs_ctx();
    if (!c) return fail(QCS_ERR_NOMEM, "context allocation failed");
    c->p = *p;
    QCS_HIP(hipGetDevice(&c->device));
    *out = c;
    return QCS_OK;
}

int qcs_ctx_set_params(qcs_ctx *c, const qcs_params *p)
{
    if (!c || !p) return fail(QCS_ERR_ARG, "null argument");
    int rc = validate_params(*p);
    if (rc) return rc;
    c->p = *p;
    return QCS_OK;
}

int qcs_ctx_destroy(qcs_ctx *c)
{
    if (!c) return QCS_OK;
    qcs::DevBuf *bufs[] = {&c->ex_bkg, &c->ex_tmp, &c->ex_det, &c->ex_mark, &c->ex_cnt, &c->ex_xy, &c->ex_var,
                           &c->ex_frames, &c->fit_idx1, &c->fit_idx2, &c->fit_idx3, &c->fit_flag, &c->fit_cnt,
                           &c->fit_rows, &c->h_rois_dev, &c->h_wle_dev, &c->h_loc_dev, &c->cf_theta, &c->cf_ll,
                           &c->cf_list, &c->cf_rows, &c->cf_queue, &c->rn_buf, &c->st_buf,
                           &c->dr_img, &c->dr_cc, &c->dr_shift, &c->dr_frames, &c->dp_hist,
                           &c->mx_rois, &c->mx_list, &c->mx_sets, &c->mx_rows};
    for (auto *b : bufs) b->release();
    c->pin_cnt.release();
    (void)timer_resolve(c->timer);
    delete[] c->timer.pend;
    delete c;
    return QCS_OK;
}

int qcs_crlb_dev(qcs_ctx *c, float *d_loc, int n, void *stream)
{
    if (!c || (!d_loc && n > 0) || n < 0) return fail(QCS_ERR_ARG, "invalid argument");
    return crlb_launch(c->p, d_loc, n, nullptr, (hipStream_t)stream);
}

int qcs_timing_enable(qcs_ctx *c, int on)
{
    if (!c) return fail(QCS_ERR_ARG, "null context");
    int rc = timer_resolve(c->timer);
    if (rc) return rc;
    for (int i = 0; i < TM_COUNT; i++) { c->timer.ms[i] = 0; c->timer.launches[i] = 0; c->timer.items[i] = 0; }
    c->timer.on = on ? 1 : 0;
    return QCS_OK;
}

int qcs_timing_get(qcs_ctx *c, qcs_timing *out)
{
    if (!c || !out) return fail(QCS_ERR_ARG, "null argument");
    int rc = timer_resolve(c->timer);
    if (rc) return rc;
    const Timer &t = c->timer;
    out->extract_ms = t.ms[TM_EXTRACT];
    out->fit_1e_ms = t.ms[TM_FIT1E];
    out->fit_me_ms = t.ms[TM_FITME];
    out->crlb_ms = t.ms[TM_CRLB];
    out->compact_ms = t.ms[TM_COMPACT];
    out->render_ms = t.ms[TM_RENDER];
    out->fit_1e_launches = t.launches[TM_FIT1E];
    out->fit_1e_rois = t.items[TM_FIT1E];
    out->fit_me_launches = t.launches[TM_FITME];
    out->cfit_ms = t.ms[TM_CFIT];
    out->nn_ms = t.ms[TM_NN];
    out->cfit_launches = t.launches[TM_CFIT];
    out->cfit_rois = t.items[TM_CFIT];
    out->nn_launches = t.launches[TM_NN];
    out->drift_cc_ms = t.ms[TM_DRIFTCC];
    out->drift_cc_launches = t.launches[TM_DRIFTCC];
    out->drift_cc_pixels = t.items[TM_DRIFTCC];
    return QCS_OK;
}

void qcs_test_fail_alloc_after(int n) { g_fail_alloc_after.store(n < 0 ? -1 : n); }

int qcs_fit_stats_get(qcs_ctx *c, qcs_fit_stats *out)
{
    if (!c || !out) return fail(QCS_ERR_ARG, "null argument");
    *out = c->stats;
    return QCS_OK;
}

}  // extern "C"
