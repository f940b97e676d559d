// session.cpp -- online localization session (the lm_* JNI flow) and the synchronous batch call.
//
// The reference's online loop (IJ/OnlineLocalizationLD.cpp:52-360, th_OnlineLocalizationLD)
// pops fed frame groups from a queue, extracts ROIs per group (frames numbered from 1), and once
// PointNumTh (20480; /4 for AS3D) ROIs have accumulated -- or the run ends -- fits them, removes
// rejected rows with frame sorting over [FirstFrame, EndFrame], merges molecules over consecutive
// frames when ConsecFitEn (the merged-away rows removed again), writes the rows to the result file,
// renders them and computes their statistics.  This file keeps that flow with HBM-resident
// accumulation: the ROI records of successive groups are appended on the device, so the only host
// traffic is the frame upload and the download of the final rows.
//
// Threading: one worker thread per session owns the context and the HIP stream; the API threads
// only touch the queue (mutex + condition variables) and read results under the same mutex.  A
// writer thread (when an output file is set) appends each flush's rows to the file, so the worker
// never waits for the disk; the session reports finished only after the file is complete.
#include <stdio.h>
#include <string.h>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <vector>
#include "qcs_internal.h"

using namespace qcs;

namespace {

constexpr int kPointNumTh = 20480;                               // bfgs_CommonPara.h:77
constexpr size_t kMaxQueuedPixels = (size_t)200 * 2048 * 2048;   // IJ/QC_STORM_CPPDLL.cpp:227

struct FeedItem {
    uint16_t *h = nullptr;   // pinned
    int nframes = 0;
    size_t bytes = 0;
};

}  // namespace

struct qcs_session {
    qcs_params p{};
    int device = 0;
    qcs_ctx *ctx = nullptr;
    hipStream_t s = nullptr;
    std::thread worker;

    std::mutex mu;                 // queue, results, flags
    std::condition_variable cv_work, cv_space;
    std::deque<FeedItem> queue;
    std::vector<FeedItem> spare;   // pinned buffers for reuse
    size_t queued_px = 0;
    bool finishing = false, finished = false, stop = false;
    int err = QCS_OK;
    std::string err_msg;
    std::vector<float> locs;       // frame-sorted rows, 12 floats each

    qcs_loc_stats stats{};         // statistics of the last flush's rows (under mu)
    float ontime[QCS_ONTIME_BINS] = {0};
    long long flushes = 0;
    long long fed = 0;             // frames accepted by qcs_session_feed (under mu)

    std::mutex dev_mu;             // device image access vs the worker's flush

    // device state (worker-owned)
    DevBuf frames, rois, wle, loc, sorted, img, zimg, merged, kept;
    HostBuf pin;                   // staging for the final rows
    qcs_consec *consec = nullptr;  // ConsecFitEn
    int acc_n = 0, first_frame = 0, end_frame = 0, cur_frame = 0, threshold = kPointNumTh, acc_cap = 0;
    long long dropped = 0;

    // result file (the reference's LocFile): chunks handed to the writer thread
    FILE *file = nullptr;
    std::thread writer;
    std::mutex wmu;
    std::condition_variable wcv;
    std::deque<std::vector<float>> wq;
    bool wdone = false;
    int werr = 0;
    std::string wpath;
};

namespace {

void set_error(qcs_session *s, int rc)
{
    std::lock_guard<std::mutex> g(s->mu);
    if (s->err == QCS_OK) {
        s->err = rc;
        s->err_msg = qcs_last_error();
    }
    s->finished = true;
    s->cv_space.notify_all();
}

int session_status(qcs_session *s)
{
    std::lock_guard<std::mutex> g(s->mu);
    if (s->err) return fail(s->err, "session worker failed: %s", s->err_msg.c_str());
    return QCS_OK;
}

void writer_main(qcs_session *s)
{
    for (;;) {
        std::vector<float> chunk;
        {
            std::unique_lock<std::mutex> g(s->wmu);
            s->wcv.wait(g, [s] { return !s->wq.empty() || s->wdone; });
            if (s->wq.empty()) return;   // done and drained
            chunk.swap(s->wq.front());
            s->wq.pop_front();
        }
        if (!s->werr && fwrite(chunk.data(), sizeof(float), chunk.size(), s->file) != chunk.size()) {
            std::lock_guard<std::mutex> g(s->wmu);
            s->werr = 1;
        }
    }
}

// ends the writer after it has written everything queued; closes the file (0 or an error code)
int close_writer(qcs_session *s)
{
    if (!s->file) return QCS_OK;
    {
        std::lock_guard<std::mutex> g(s->wmu);
        s->wdone = true;
        s->wcv.notify_all();
    }
    if (s->writer.joinable()) s->writer.join();
    const bool bad = s->werr || fclose(s->file) != 0;
    s->file = nullptr;
    if (bad) return fail(QCS_ERR_IO, "writing the result file %s failed", s->wpath.c_str());
    return QCS_OK;
}

// Fit the accumulated ROIs, compact/sort by frame, merge consecutive frames, write, render, compute
// statistics and append the rows.  is_end: the last flush (gives back the rows the merger holds).
int flush(qcs_session *s, bool is_end)
{
    const bool consec = s->consec != nullptr;
    if (s->acc_n == 0 && !(is_end && consec && qcs_consec_held(s->consec) > 0)) return QCS_OK;
    const int n = s->acc_n;
    std::lock_guard<std::mutex> dg(s->dev_mu);
    int rows = 0, m = 0, rc;
    if (n > 0) {
        if ((rc = qcs_fit_dev(s->ctx, s->rois.as<uint16_t>(), s->wle.as<float>(), n, s->loc.as<float>(), 2 * s->acc_cap,
                              &rows, s->s)))
            return rc;
        if ((rc = qcs_compact_sorted_dev(s->ctx, s->loc.as<float>(), rows, s->first_frame, s->end_frame,
                                         s->sorted.as<float>(), &m, s->s)))
            return rc;
    }
    const float *res = s->sorted.as<float>();
    float ontime[QCS_ONTIME_BINS] = {0};
    if (consec) {
        // ConsecutiveFit_WeightedAvg, then RemoveZeroLocalizations without sorting
        // (IJ/OnlineLocalizationLD.cpp:240-256)
        const int held = qcs_consec_held(s->consec);
        const size_t cap = (size_t)held + (size_t)m;
        const size_t rb = (size_t)QCS_REC_FLOATS * sizeof(float);
        if ((rc = s->merged.ensure((cap > 0 ? cap : 1) * rb)) || (rc = s->kept.ensure((cap > 0 ? cap : 1) * rb))) return rc;
        int k = 0;
        if ((rc = qcs_consec_fit_dev(s->consec, s->sorted.as<float>(), m, is_end ? 1 : 0, s->merged.as<float>(), (int)cap, &k,
                                     ontime, s->s)) ||
            (rc = qcs_compact_dev(s->ctx, s->merged.as<float>(), k, s->kept.as<float>(), &m, s->s)))
            return rc;
        res = s->kept.as<float>();
    }
    qcs_loc_stats st;
    if ((rc = qcs_loc_stats_dev(s->ctx, res, m, &st, s->s))) return rc;   // GetStatisticalInf per flush
    if (m > 0) {
        if ((rc = qcs_render_dev(s->ctx, res, m, s->img.as<float>(), s->zimg.as<float>(), s->s))) return rc;
        const size_t b = (size_t)m * QCS_REC_FLOATS * sizeof(float);
        if ((rc = s->pin.ensure(b))) return rc;
        QCS_HIP(hipMemcpyAsync(s->pin.ptr, res, b, hipMemcpyDeviceToHost, s->s));
    }
    QCS_HIP(hipStreamSynchronize(s->s));
    const float *src = s->pin.as<float>();
    if (m > 0 && s->file) {
        std::lock_guard<std::mutex> g(s->wmu);
        s->wq.emplace_back(src, src + (size_t)m * QCS_REC_FLOATS);
        s->wcv.notify_one();
    }
    {
        std::lock_guard<std::mutex> g(s->mu);
        if (m > 0) s->locs.insert(s->locs.end(), src, src + (size_t)m * QCS_REC_FLOATS);
        s->stats = st;
        if (consec) memcpy(s->ontime, ontime, sizeof(ontime));
        s->flushes++;
    }
    s->acc_n = 0;
    return QCS_OK;
}

// Extract one fed group (its frames form one batched image, as in ExtractMolecules).
int process(qcs_session *s, const FeedItem &it)
{
    const size_t fb = it.bytes;
    int rc;
    if ((rc = s->frames.ensure(fb))) return rc;
    QCS_HIP(hipMemcpyAsync(s->frames.ptr, it.h, fb, hipMemcpyHostToDevice, s->s));
    s->ctx->p.BatchedImageNum = it.nframes;
    const int R = s->p.ROISize;
    int no = 0, nf = 0;
    rc = qcs_extract_dev(s->ctx, s->frames.as<uint16_t>(), it.nframes, s->cur_frame + 1,
                         s->rois.as<uint16_t>() + (size_t)s->acc_n * QCS_ROI_WORDS(R),
                         s->wle.as<float>() + (size_t)s->acc_n * QCS_WLE_FLOATS, s->acc_cap - s->acc_n, &no, &nf, s->s);
    if (rc && rc != QCS_ERR_CAPACITY) return rc;
    s->dropped += nf - no;   // beyond MaxROIPerBatch: the reference drops them too
    if (s->acc_n == 0) s->first_frame = s->cur_frame + 1;
    s->cur_frame += it.nframes;
    s->end_frame = s->cur_frame;
    s->acc_n += no;
    if (s->acc_n >= s->threshold) return flush(s, false);
    return QCS_OK;
}

void worker_main(qcs_session *s)
{
    if (hipSetDevice(s->device) != hipSuccess) {
        fail(QCS_ERR_HIP, "hipSetDevice(%d) failed in the session worker", s->device);
        set_error(s, QCS_ERR_HIP);
        return;
    }
    for (;;) {
        FeedItem it;
        {
            std::unique_lock<std::mutex> g(s->mu);
            s->cv_work.wait(g, [s] { return !s->queue.empty() || s->finishing || s->stop; });
            if (s->stop) return;
            if (s->queue.empty()) break;   // finishing and drained
            it = s->queue.front();
            s->queue.pop_front();
        }
        int rc = process(s, it);
        {
            std::lock_guard<std::mutex> g(s->mu);
            s->queued_px -= (size_t)it.nframes * s->p.ImageWidth * s->p.ImageHigh;
            s->spare.push_back(it);
            s->cv_space.notify_all();
        }
        if (rc) {
            set_error(s, rc);
            return;
        }
    }
    int rc = flush(s, true);
    if (!rc) rc = close_writer(s);   // the file is complete before the session reports finished
    if (rc) {
        set_error(s, rc);
        return;
    }
    std::lock_guard<std::mutex> g(s->mu);
    s->finished = true;
    s->cv_space.notify_all();
}

void free_items(std::vector<FeedItem> &v)
{
    for (auto &it : v) (void)hipHostFree(it.h);
    v.clear();
}

}  // namespace

extern "C" {

int qcs_session_create(const qcs_params *pin, qcs_session **out)
{
    if (!pin || !out) return fail(QCS_ERR_ARG, "null argument");
    *out = nullptr;
    qcs_params p = *pin;
    int rc = validate_params(p);
    if (rc) return rc;
    if (p.ImageWidth <= 0 || p.ImageHigh <= 0) return fail(QCS_ERR_ARG, "ImageWidth/ImageHigh not set");
    if (p.LocType == QCS_LOCTYPE_DH3D) return fail(QCS_ERR_ARG, "LocType DH3D is not supported");
    if (p.MultiEmitterFitEn && p.ROISize > 13)
        return fail(QCS_ERR_ARG, "multi-emitter fitting supports ROISize 5..13 (got %d)", p.ROISize);
    if (p.SRImageWidth <= 16 || p.SRImageHigh <= 16) qcs_params_update_sr_size(&p);
    qcs_session *s = new (std::nothrow) qcs_session();
    if (!s) return fail(QCS_ERR_NOMEM, "session allocation failed");
    s->p = p;
    s->threshold = p.FlushROINum > 0 ? p.FlushROINum : (p.LocType == QCS_LOCTYPE_AS3D ? kPointNumTh / 4 : kPointNumTh);
    const int batch_cap = p.MaxROIPerBatch > 0 ? p.MaxROIPerBatch : QCS_MAX_POINT_NUM;
    s->acc_cap = s->threshold + batch_cap;
    const int R = p.ROISize;
    const size_t npx = (size_t)p.SRImageWidth * p.SRImageHigh;
    auto cleanup = [s](int code) {
        if (s->consec) qcs_consec_destroy(s->consec);
        if (s->ctx) qcs_ctx_destroy(s->ctx);
        if (s->s) (void)hipStreamDestroy(s->s);
        DevBuf *b[] = {&s->frames, &s->rois, &s->wle, &s->loc, &s->sorted, &s->img, &s->zimg};
        for (auto *x : b) x->release();
        s->pin.release();
        delete s;
        return code;
    };
    if (hipGetDevice(&s->device) != hipSuccess) return cleanup(fail(QCS_ERR_NODEVICE, "no HIP device"));
    if ((rc = qcs_ctx_create(&p, &s->ctx))) return cleanup(rc);
    if (p.ConsecFitEn && (rc = qcs_consec_create(s->ctx, &s->consec))) return cleanup(rc);
    if (hipStreamCreateWithFlags(&s->s, hipStreamNonBlocking) != hipSuccess)
        return cleanup(fail(QCS_ERR_HIP, "hipStreamCreate failed"));
    if ((rc = s->rois.ensure((size_t)s->acc_cap * QCS_ROI_WORDS(R) * sizeof(uint16_t))) ||
        (rc = s->wle.ensure((size_t)s->acc_cap * QCS_WLE_FLOATS * sizeof(float))) ||
        (rc = s->loc.ensure((size_t)2 * s->acc_cap * QCS_REC_FLOATS * sizeof(float))) ||
        (rc = s->sorted.ensure((size_t)2 * s->acc_cap * QCS_REC_FLOATS * sizeof(float))) ||
        (rc = s->img.ensure(npx * sizeof(float))) ||
        (p.LocType == QCS_LOCTYPE_AS3D && (rc = s->zimg.ensure(npx * sizeof(float)))))
        return cleanup(rc);
    if (hipMemsetAsync(s->img.ptr, 0, npx * sizeof(float), s->s) != hipSuccess ||
        (s->zimg.ptr && hipMemsetAsync(s->zimg.ptr, 0, npx * sizeof(float), s->s) != hipSuccess) ||
        hipStreamSynchronize(s->s) != hipSuccess)
        return cleanup(fail(QCS_ERR_HIP, "SR image clear failed"));
    try {
        s->worker = std::thread(worker_main, s);
    } catch (...) {
        return cleanup(fail(QCS_ERR_NOMEM, "could not start the session worker"));
    }
    *out = s;
    return QCS_OK;
}

int qcs_session_feed(qcs_session *s, const uint16_t *frames, int nframes)
{
    if (!s || nframes < 0 || (nframes > 0 && !frames)) return fail(QCS_ERR_ARG, "invalid argument");
    if (nframes == 0) return session_status(s);
    const size_t px = (size_t)nframes * s->p.ImageWidth * s->p.ImageHigh;
    const size_t bytes = px * sizeof(uint16_t);
    FeedItem it;
    {
        std::unique_lock<std::mutex> g(s->mu);
        if (s->finishing) return fail(QCS_ERR_STATE, "session already finished");
        // back-pressure: wait while the queue holds the reference's buffer budget
        s->cv_space.wait(g, [&] { return s->queued_px == 0 || s->queued_px + px <= kMaxQueuedPixels || s->err; });
        if (s->err) return fail(s->err, "session worker failed: %s", s->err_msg.c_str());
        for (size_t i = 0; i < s->spare.size(); i++)
            if (s->spare[i].bytes == bytes) {
                it = s->spare[i];
                s->spare.erase(s->spare.begin() + (long)i);
                break;
            }
        s->queued_px += px;   // reserve before the copy so concurrent feeders see it
    }
    if (!it.h) {
        void *h = nullptr;
        if (hipHostMalloc(&h, bytes, hipHostMallocDefault) != hipSuccess) {
            std::lock_guard<std::mutex> g(s->mu);
            s->queued_px -= px;
            return fail(QCS_ERR_NOMEM, "pinned allocation of %zu bytes failed", bytes);
        }
        it.h = static_cast<uint16_t *>(h);
        it.bytes = bytes;
    }
    memcpy(it.h, frames, bytes);
    it.nframes = nframes;
    std::lock_guard<std::mutex> g(s->mu);
    s->fed += nframes;
    s->queue.push_back(it);
    s->cv_work.notify_one();
    return QCS_OK;
}

int qcs_session_finish(qcs_session *s)
{
    if (!s) return fail(QCS_ERR_ARG, "null session");
    std::lock_guard<std::mutex> g(s->mu);
    s->finishing = true;
    s->cv_work.notify_one();
    if (s->err) return fail(s->err, "session worker failed: %s", s->err_msg.c_str());
    return QCS_OK;
}

int qcs_session_is_finished(qcs_session *s)
{
    if (!s) return 0;
    std::lock_guard<std::mutex> g(s->mu);
    return s->finished ? 1 : 0;
}

long long qcs_session_fetch_locs(qcs_session *s, float *out12, long long cap)
{
    if (!s || cap < 0) return -fail(QCS_ERR_ARG, "invalid argument");
    std::lock_guard<std::mutex> g(s->mu);
    if (s->err) return -fail(s->err, "session worker failed: %s", s->err_msg.c_str());
    const long long n = (long long)(s->locs.size() / QCS_REC_FLOATS);
    if (out12 && cap > 0) {
        const long long k = n < cap ? n : cap;
        memcpy(out12, s->locs.data(), (size_t)k * QCS_REC_FLOATS * sizeof(float));
    }
    return n;
}

int qcs_session_get_sr_image(qcs_session *s, float *img)
{
    if (!s || !img) return fail(QCS_ERR_ARG, "invalid argument");
    int rc = session_status(s);
    if (rc) return rc;
    std::lock_guard<std::mutex> dg(s->dev_mu);
    QCS_HIP(hipSetDevice(s->device));
    QCS_HIP(hipStreamSynchronize(s->s));
    QCS_HIP(hipMemcpy(img, s->img.ptr, (size_t)s->p.SRImageWidth * s->p.SRImageHigh * sizeof(float),
                      hipMemcpyDeviceToHost));
    return QCS_OK;
}

int qcs_session_get_sr_image3d(qcs_session *s, uint8_t *bgra)
{
    if (!s || !bgra) return fail(QCS_ERR_ARG, "invalid argument");
    if (s->p.LocType != QCS_LOCTYPE_AS3D || !s->zimg.ptr) return fail(QCS_ERR_STATE, "the 3D image needs an AS3D session");
    int rc = session_status(s);
    if (rc) return rc;
    std::lock_guard<std::mutex> dg(s->dev_mu);
    QCS_HIP(hipSetDevice(s->device));
    const size_t npx = (size_t)s->p.SRImageWidth * s->p.SRImageHigh;
    DevBuf d;
    if ((rc = d.ensure(npx * 4))) return rc;
    rc = qcs_render_rgb_dev(s->ctx, s->img.as<float>(), s->zimg.as<float>(), QCS_RGB_BGRA, 1.0f, d.as<uint8_t>(), nullptr, s->s);
    hipError_t e = hipSuccess;
    if (!rc) e = hipMemcpyAsync(bgra, d.ptr, npx * 4, hipMemcpyDeviceToHost, s->s);
    if (!rc && e == hipSuccess) e = hipStreamSynchronize(s->s);
    d.release();
    if (rc) return rc;
    if (e != hipSuccess) return hip_fail(e, "3D image", __FILE__, __LINE__);
    return QCS_OK;
}

int qcs_session_get_max_disp_val(qcs_session *s, int *val)
{
    if (!s || !val) return fail(QCS_ERR_ARG, "invalid argument");
    int rc = session_status(s);
    if (rc) return rc;
    std::lock_guard<std::mutex> dg(s->dev_mu);
    QCS_HIP(hipSetDevice(s->device));
    DevBuf d;
    if ((rc = d.ensure(sizeof(int)))) return rc;
    int h = 0;
    hipError_t e = hipMemsetAsync(d.ptr, 0, sizeof(int), s->s);
    if (e == hipSuccess) {
        rc = max_disp_launch(s->img.as<float>(), (size_t)s->p.SRImageWidth * s->p.SRImageHigh, d.as<int>(), s->s);
        if (!rc) e = hipMemcpyAsync(&h, d.ptr, sizeof(int), hipMemcpyDeviceToHost, s->s);
        if (!rc && e == hipSuccess) e = hipStreamSynchronize(s->s);
    }
    d.release();
    if (rc) return rc;
    if (e != hipSuccess) return hip_fail(e, "max display value", __FILE__, __LINE__);
    *val = h > 1 ? h : 1;   // the reference's DispMaxVal starts at 1 (IJ/QC_STORM_CPPDLL.cpp:194-205)
    return QCS_OK;
}

int qcs_session_destroy(qcs_session *s)
{
    if (!s) return QCS_OK;
    {
        std::lock_guard<std::mutex> g(s->mu);
        s->finishing = true;
        s->stop = true;   // lm_ReleaseResource: queued frames not yet processed are discarded
        s->cv_work.notify_all();
    }
    if (s->worker.joinable()) s->worker.join();
    (void)close_writer(s);
    (void)hipSetDevice(s->device);
    for (auto &it : s->queue) (void)hipHostFree(it.h);
    s->queue.clear();
    free_items(s->spare);
    if (s->s) (void)hipStreamSynchronize(s->s);
    DevBuf *b[] = {&s->frames, &s->rois, &s->wle, &s->loc, &s->sorted, &s->img, &s->zimg, &s->merged, &s->kept};
    for (auto *x : b) x->release();
    s->pin.release();
    if (s->consec) qcs_consec_destroy(s->consec);
    if (s->ctx) qcs_ctx_destroy(s->ctx);
    if (s->s) (void)hipStreamDestroy(s->s);
    delete s;
    return QCS_OK;
}

int qcs_session_set_output(qcs_session *s, const char *path)
{
    if (!s || !path || !*path) return fail(QCS_ERR_ARG, "invalid argument");
    {
        std::lock_guard<std::mutex> g(s->mu);
        if (s->finishing || s->fed > 0)
            return fail(QCS_ERR_STATE, "the result file must be set before the first frames are fed");
    }
    if (s->file) return fail(QCS_ERR_STATE, "the result file is already set");
    FILE *f = fopen(path, "wb");
    if (!f) return fail(QCS_ERR_IO, "cannot create the result file %s", path);
    s->wpath = path;
    s->file = f;
    try {
        s->writer = std::thread(writer_main, s);
    } catch (...) {
        fclose(f);
        s->file = nullptr;
        return fail(QCS_ERR_NOMEM, "could not start the result writer");
    }
    return QCS_OK;
}

int qcs_session_get_stats(qcs_session *s, qcs_loc_stats *out, float *ontime_ratio)
{
    if (!s || (!out && !ontime_ratio)) return fail(QCS_ERR_ARG, "invalid argument");
    std::lock_guard<std::mutex> g(s->mu);
    if (s->err) return fail(s->err, "session worker failed: %s", s->err_msg.c_str());
    if (out) *out = s->stats;
    if (ontime_ratio) memcpy(ontime_ratio, s->ontime, sizeof(s->ontime));
    return QCS_OK;
}

int qcs_result_file_name(const qcs_params *p, const char *image_name, char *out, int cap)
{
    if (!p || !image_name || cap < 0 || (cap > 0 && !out)) return -fail(QCS_ERR_ARG, "invalid argument");
    // SetLocDataFileName (IJ/LocResource.cpp:160-179): CString::TrimRight(".tif") drops every
    // trailing character found in the set {'.', 't', 'i', 'f'}, then the suffix is appended
    std::string base = image_name;
    while (!base.empty() && strchr(".tif", base.back()) != nullptr) base.pop_back();
    char consec[64] = "";
    if (p->ConsecFitEn) snprintf(consec, sizeof(consec), "_Consec%.0fnm", p->ConsecFit_DistanceTh_nm);
    char suffix[128];
    snprintf(suffix, sizeof(suffix), "_result%dD%d%s%s.txt", p->LocType + 2, p->ROISize, p->MultiEmitterFitEn ? "_M" : "_S",
             consec);
    const std::string name = base + suffix;
    if (out && cap > 0) {
        const size_t k = name.size() < (size_t)cap - 1 ? name.size() : (size_t)cap - 1;
        memcpy(out, name.data(), k);
        out[k] = 0;
    }
    return (int)name.size();
}

long long qcs_loc_batch_sync(qcs_ctx *c, const uint16_t *frames, int nframes, int start_frame, float *out12,
                             long long cap)
{
    return qcs_loc_batch_sync_info(c, frames, nframes, start_frame, out12, cap, nullptr, nullptr);
}

long long qcs_loc_batch_sync_info(qcs_ctx *c, const uint16_t *frames, int nframes, int start_frame, float *out12,
                                  long long cap, float *info8, qcs_loc_stats *stats)
{
    if (!c || nframes < 0 || cap < 0 || (nframes > 0 && !frames)) return -fail(QCS_ERR_ARG, "invalid argument");
    if (info8) memset(info8, 0, 8 * sizeof(float));
    if (stats) memset(stats, 0, sizeof(*stats));
    if (nframes == 0) return 0;
    const qcs_params saved = c->p;
    if (c->p.ImageWidth <= 0 || c->p.ImageHigh <= 0) return -fail(QCS_ERR_ARG, "ImageWidth/ImageHigh not set");
    const int R = c->p.ROISize;
    const int rcap = c->p.MaxROIPerBatch > 0 ? c->p.MaxROIPerBatch : QCS_MAX_POINT_NUM;
    const size_t fb = (size_t)nframes * c->p.ImageWidth * c->p.ImageHigh * sizeof(uint16_t);
    int rc;
    if ((rc = c->ex_frames.ensure(fb)) || (rc = c->h_rois_dev.ensure((size_t)rcap * QCS_ROI_WORDS(R) * 2)) ||
        (rc = c->h_wle_dev.ensure((size_t)rcap * QCS_WLE_FLOATS * 4)) ||
        (rc = c->h_loc_dev.ensure((size_t)4 * rcap * QCS_REC_FLOATS * 4)))
        return -rc;
    if (hipMemcpy(c->ex_frames.ptr, frames, fb, hipMemcpyHostToDevice) != hipSuccess)
        return -fail(QCS_ERR_HIP, "frame upload failed");
    c->p.BatchedImageNum = nframes;   // the call's frames form one batched image (lm_LocBatchedImg)
    int no = 0, nf = 0, rows = 0, m = 0;
    float *loc = c->h_loc_dev.as<float>(), *out = loc + (size_t)2 * rcap * QCS_REC_FLOATS;
    rc = qcs_extract_dev(c, c->ex_frames.as<uint16_t>(), nframes, start_frame, c->h_rois_dev.as<uint16_t>(),
                         c->h_wle_dev.as<float>(), rcap, &no, &nf, nullptr);
    c->p = saved;
    if (rc && rc != QCS_ERR_CAPACITY) return -rc;
    if ((rc = qcs_fit_dev(c, c->h_rois_dev.as<uint16_t>(), c->h_wle_dev.as<float>(), no, loc, 2 * rcap, &rows,
                          nullptr)))
        return -rc;
    if ((rc = qcs_compact_sorted_dev(c, loc, rows, start_frame, start_frame + nframes - 1, out, &m, nullptr))) return -rc;
    if (out12 && cap > 0 && m > 0) {
        const long long k = m < cap ? m : cap;
        if (hipMemcpy(out12, out, (size_t)k * QCS_REC_FLOATS * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
            return -fail(QCS_ERR_HIP, "result download failed");
    }
    if (info8 || stats) {
        // statistics of the batch's rows and the MM summary (MM/QC_STORM_CPPDLL.cpp:895-913)
        qcs_loc_stats st;
        if ((rc = qcs_loc_stats_dev(c, out, m, &st, nullptr))) return -rc;
        if (stats) *stats = st;
        if (info8) {
            info8[0] = st.mean_peak_snr;
            info8[1] = -st.mean_psf_width;
            info8[2] = st.mean_peak_snr - st.mean_psf_width;
            info8[3] = st.mean_peak_snr * (-st.mean_psf_width);
            info8[4] = (float)m;
            info8[5] = st.loc_density_2d;
        }
    }
    return m;
}

}  // extern "C"
