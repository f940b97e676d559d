// qcstorm_jni.cpp -- JNI forwarding layer for the QC-STORM ImageJ and Micro-Manager plug-ins.
//
// The Java classes keep their native declarations (QC_STORM_ for ImageJ, IJJ/QC_STORM_.java:79-117;
// hust.whno.SMLM.QC_STORM_Plug for Micro-Manager, MMJ/QC_STORM_Plug.java:41-94); this file
// defines the natives they bind and forwards each to libqcstorm's C-ABI (include/qcstorm.h).  It
// holds the plug-in state the reference keeps in globals: the run parameters
// (LocPara_Global), the online session, the image name / save path that name the result file,
// and, for Micro-Manager's synchronous batch call, one context.
//
// Reference natives replaced (IJ = Plugin for ImageJ/QC-STORM_ CPPDLL/QC-STORM_ CPPDLL,
// MM = Plugin for Micro-Manager/QC-STORM CPPDLL/QC-STORM CPPDLL):
//   IJ/QC_STORM_CPPDLL.h:15-174 (definitions IJ/QC_STORM_CPPDLL.cpp:33-460)
//   MM/QC_STORM_CPPDLL.cpp:38-924
// Natives of the feedback, stage, laser, statistics-image, re-render and spatial-resolution
// displays are outside the localization path and are not defined here.
//
// Built only where a JDK provides jni.h: make -C qc-storm_amd jni JAVA_HOME=/path/to/jdk
#include <jni.h>
#include <stdio.h>
#include <string.h>
#include <mutex>
#include <string>
#include <vector>
#include "../../include/qcstorm.h"

namespace {

struct PluginState {
    std::mutex mu;
    qcs_params para;
    bool para_init = false;
    qcs_session *sess = nullptr;
    qcs_ctx *ctx = nullptr;            // Micro-Manager lm_LocBatchedImg
    std::string image_name;            // ImageJ: the result file is named after the image
    std::string save_path, create_time;  // Micro-Manager: <save path>loc_result..._<time>...
    std::vector<float> rows;           // Micro-Manager lm_LocBatchedImg row buffer
};

PluginState &state()
{
    static PluginState s;
    return s;
}

void para_init(PluginState &s)
{
    if (!s.para_init) {
        qcs_params_default(&s.para);
        s.para_init = true;
    }
}

void report(const char *what, int rc)
{
    if (rc != QCS_OK) fprintf(stderr, "qcstorm: %s failed (%d): %s\n", what, rc, qcs_last_error());
}

std::string jstr(JNIEnv *env, jstring js)
{
    if (!js) return std::string();
    const char *c = env->GetStringUTFChars(js, nullptr);
    std::string r = c ? c : "";
    if (c) env->ReleaseStringUTFChars(js, c);
    return r;
}

void set_loc_para(PluginState &s, jfloat kadc, jfloat offset, jfloat qe, jint roi, jint loctype, jint multi, jint wle,
                  jint consec, jfloat consec_nm, jfloat px, jfloat zoom, jfloat snr)
{
    para_init(s);
    s.para.KAdc = kadc;
    s.para.Offset = offset;
    s.para.QE = qe;
    s.para.ROISize = roi;
    s.para.LocType = loctype;
    s.para.MultiEmitterFitEn = multi;
    s.para.WLEEn = wle;
    s.para.ConsecFitEn = consec;
    s.para.ConsecFit_DistanceTh_nm = consec_nm;
    s.para.PixelSize = px;
    s.para.PixelZoom = zoom;
    s.para.SNR_th = snr;
}

void set_loc_para_3d(PluginState &s, jfloat minz, jfloat maxz, jfloat zcorr, jfloat g4, jfloat g3, jfloat g2, jfloat g1,
                     jfloat g0, jfloat l4, jfloat l3, jfloat l2, jfloat l1, jfloat l0)
{
    para_init(s);
    s.para.MinZDepth = minz;
    s.para.MaxZDepth = maxz;
    s.para.ZDepthCorrFactor = zcorr;
    s.para.p4_XGY = g4; s.para.p3_XGY = g3; s.para.p2_XGY = g2; s.para.p1_XGY = g1; s.para.p0_XGY = g0;
    s.para.p4_XLY = l4; s.para.p3_XLY = l3; s.para.p2_XLY = l2; s.para.p1_XLY = l1; s.para.p0_XLY = l0;
}

// the online session with its result file (IJ/OnlineLocalizationLD.cpp:70-71, MM/OnlineLocalizationLD.cpp:107-118)
void start_session(PluginState &s, bool micro_manager)
{
    if (s.sess) return;
    para_init(s);
    int rc = qcs_session_create(&s.para, &s.sess);
    if (rc != QCS_OK) {
        report("qcs_session_create", rc);
        s.sess = nullptr;
        return;
    }
    std::string file;
    if (micro_manager) {
        char tail[256];
        if (s.para.ConsecFitEn)
            snprintf(tail, sizeof(tail), "%s_Consec%.0fnm", s.para.MultiEmitterFitEn ? "_M" : "_S", s.para.ConsecFit_DistanceTh_nm);
        else
            snprintf(tail, sizeof(tail), "%s", s.para.MultiEmitterFitEn ? "_M" : "_S");
        char name[512];
        snprintf(name, sizeof(name), "loc_result%dD%d_%s%s.txt", s.para.LocType + 2, s.para.ROISize, s.create_time.c_str(), tail);
        file = s.save_path + name;
    } else {
        const int len = qcs_result_file_name(&s.para, s.image_name.c_str(), nullptr, 0);
        if (len > 0) {
            std::vector<char> buf((size_t)len + 1);
            qcs_result_file_name(&s.para, s.image_name.c_str(), buf.data(), len + 1);
            file = buf.data();
        }
    }
    if (!file.empty()) report("qcs_session_set_output", qcs_session_set_output(s.sess, file.c_str()));
}

void feed(PluginState &s, JNIEnv *env, jshortArray a, jint nframes)
{
    if (!s.sess || nframes <= 0) return;
    jshort *p = (jshort *)env->GetPrimitiveArrayCritical(a, nullptr);
    if (!p) return;
    const int rc = qcs_session_feed(s.sess, (const uint16_t *)p, nframes);   // copies before returning
    env->ReleasePrimitiveArrayCritical(a, p, JNI_ABORT);
    report("qcs_session_feed", rc);
}

jfloatArray sr_image(PluginState &s, JNIEnv *env)
{
    const jsize n = (jsize)s.para.SRImageWidth * s.para.SRImageHigh;
    std::vector<float> img((size_t)(n > 0 ? n : 0));
    if (s.sess && n > 0) report("qcs_session_get_sr_image", qcs_session_get_sr_image(s.sess, img.data()));
    jfloatArray r = env->NewFloatArray(n);
    if (r && n > 0) env->SetFloatArrayRegion(r, 0, n, img.data());
    return r;
}

jintArray sr_image_3d(PluginState &s, JNIEnv *env)
{
    const jsize n = (jsize)s.para.SRImageWidth * s.para.SRImageHigh;
    std::vector<jint> img((size_t)(n > 0 ? n : 0));   // 4 bytes per pixel: B, G, R, A
    if (s.sess && n > 0) report("qcs_session_get_sr_image3d", qcs_session_get_sr_image3d(s.sess, (uint8_t *)img.data()));
    jintArray r = env->NewIntArray(n);
    if (r && n > 0) env->SetIntArrayRegion(r, 0, n, img.data());
    return r;
}

jint max_disp(PluginState &s)
{
    int v = 1;
    if (s.sess) qcs_session_get_max_disp_val(s.sess, &v);
    return v;
}

void release(PluginState &s)
{
    if (s.sess) {
        qcs_session_destroy(s.sess);   // completes the result file
        s.sess = nullptr;
    }
    if (s.ctx) {
        qcs_ctx_destroy(s.ctx);
        s.ctx = nullptr;
    }
}

}  // namespace

extern "C" {

// ---- ImageJ: class QC_STORM_ (IJ/QC_STORM_CPPDLL.h:15-174) ------------------------------------
JNIEXPORT void JNICALL Java_QC_1STORM_1_lm_1SetImagePara(JNIEnv *env, jclass, jint w, jint h, jint srw, jint srh, jint nframes,
                                                         jstring name)
{
    PluginState &s = state();
    std::lock_guard<std::mutex> lk(s.mu);
    para_init(s);
    s.para.ImageWidth = w;
    s.para.ImageHigh = h;
    s.para.SRImageWidth = srw;
    s.para.SRImageHigh = srh;
    s.para.TotalFrameNum = nframes;
    s.image_name = jstr(env, name);
}

JNIEXPORT void JNICALL Java_QC_1STORM_1_lm_1SetLocPara(JNIEnv *, jclass, jfloat kadc, jfloat offset, jfloat qe, jint roi,
                                                       jint loctype, jint multi, jint wle, jint consec, jfloat consec_nm,
                                                       jfloat px, jfloat zoom, jfloat snr)
{
    PluginState &s = state();
    std::lock_guard<std::mutex> lk(s.mu);
    set_loc_para(s, kadc, offset, qe, roi, loctype, multi, wle, consec, consec_nm, px, zoom, snr);
}

JNIEXPORT void JNICALL Java_QC_1STORM_1_lm_1SetLocPara3D(JNIEnv *, jclass, jfloat minz, jfloat maxz, jfloat zcorr, jfloat g4,
                                                         jfloat g3, jfloat g2, jfloat g1, jfloat g0, jfloat l4, jfloat l3,
                                                         jfloat l2, jfloat l1, jfloat l0)
{
    PluginState &s = state();
    std::lock_guard<std::mutex> lk(s.mu);
    set_loc_para_3d(s, minz, maxz, zcorr, g4, g3, g2, g1, g0, l4, l3, l2, l1, l0);
}

JNIEXPORT void JNICALL Java_QC_1STORM_1_lm_1StartLocThread(JNIEnv *, jclass)
{
    PluginState &s = state();
    std::lock_guard<std::mutex> lk(s.mu);
    start_session(s, false);
}

JNIEXPORT void JNICALL Java_QC_1STORM_1_lm_1FeedImageData(JNIEnv *env, jclass, jshortArray a, jint n)
{
    PluginState &s = state();
    std::lock_guard<std::mutex> lk(s.mu);
    feed(s, env, a, n);
}

JNIEXPORT void JNICALL Java_QC_1STORM_1_lm_1StopLocThread(JNIEnv *, jclass)
{
    PluginState &s = state();
    std::lock_guard<std::mutex> lk(s.mu);
    if (s.sess) report("qcs_session_finish", qcs_session_finish(s.sess));
}

JNIEXPORT jint JNICALL Java_QC_1STORM_1_lm_1IsLocFinish(JNIEnv *, jclass)
{
    PluginState &s = state();
    std::lock_guard<std::mutex> lk(s.mu);
    return s.sess ? qcs_session_is_finished(s.sess) : 1;
}

JNIEXPORT jint JNICALL Java_QC_1STORM_1_lm_1GetMaxDispVal(JNIEnv *, jclass)
{
    PluginState &s = state();
    std::lock_guard<std::mutex> lk(s.mu);
    return max_disp(s);
}

JNIEXPORT jfloatArray JNICALL Java_QC_1STORM_1_lm_1GetSMLMImage(JNIEnv *env, jclass)
{
    PluginState &s = state();
    std::lock_guard<std::mutex> lk(s.mu);
    return sr_image(s, env);
}

JNIEXPORT jintArray JNICALL Java_QC_1STORM_1_lm_1GetSMLMImage3D(JNIEnv *env, jclass)
{
    PluginState &s = state();
    std::lock_guard<std::mutex> lk(s.mu);
    return sr_image_3d(s, env);
}

JNIEXPORT void JNICALL Java_QC_1STORM_1_lm_1ReleaseResource(JNIEnv *, jclass)
{
    PluginState &s = state();
    std::lock_guard<std::mutex> lk(s.mu);
    release(s);
}

// ---- Micro-Manager: class hust.whno.SMLM.QC_STORM_Plug (MM/QC_STORM_CPPDLL.cpp:38-924) ---------
JNIEXPORT void JNICALL Java_hust_whno_SMLM_QC_1STORM_1Plug_lm_1SetSavePath(JNIEnv *env, jclass, jstring path)
{
    PluginState &s = state();
    std::lock_guard<std::mutex> lk(s.mu);
    s.save_path = jstr(env, path);
}

JNIEXPORT void JNICALL Java_hust_whno_SMLM_QC_1STORM_1Plug_lm_1SetAcquisitionPara(JNIEnv *env, jclass, jstring create_time)
{
    PluginState &s = state();
    std::lock_guard<std::mutex> lk(s.mu);
    s.create_time = jstr(env, create_time);
}

JNIEXPORT void JNICALL Java_hust_whno_SMLM_QC_1STORM_1Plug_lm_1SetLocPara(JNIEnv *, jclass, jfloat kadc, jfloat offset, jfloat qe,
                                                                          jint roi, jint loctype, jint multi, jint wle,
                                                                          jint consec, jfloat consec_nm, jfloat px,
                                                                          jfloat zoom, jfloat snr)
{
    PluginState &s = state();
    std::lock_guard<std::mutex> lk(s.mu);
    set_loc_para(s, kadc, offset, qe, roi, loctype, multi, wle, consec, consec_nm, px, zoom, snr);
}

JNIEXPORT void JNICALL Java_hust_whno_SMLM_QC_1STORM_1Plug_lm_1SetLocPara3D(JNIEnv *, jclass, jfloat minz, jfloat maxz,
                                                                            jfloat zcorr, jfloat g4, jfloat g3, jfloat g2,
                                                                            jfloat g1, jfloat g0, jfloat l4, jfloat l3,
                                                                            jfloat l2, jfloat l1, jfloat l0)
{
    PluginState &s = state();
    std::lock_guard<std::mutex> lk(s.mu);
    set_loc_para_3d(s, minz, maxz, zcorr, g4, g3, g2, g1, g0, l4, l3, l2, l1, l0);
}

JNIEXPORT void JNICALL Java_hust_whno_SMLM_QC_1STORM_1Plug_lm_1SetImagePara(JNIEnv *, jclass, jint w, jint h, jint srw, jint srh)
{
    PluginState &s = state();
    std::lock_guard<std::mutex> lk(s.mu);
    para_init(s);
    s.para.ImageWidth = w;
    s.para.ImageHigh = h;
    s.para.SRImageWidth = srw;
    s.para.SRImageHigh = srh;
}

JNIEXPORT void JNICALL Java_hust_whno_SMLM_QC_1STORM_1Plug_lm_1StartLocThread(JNIEnv *, jclass)
{
    PluginState &s = state();
    std::lock_guard<std::mutex> lk(s.mu);
    start_session(s, true);
}

// one frame per call; the first feed starts the session (MM/QC_STORM_CPPDLL.cpp:224, 263-268)
JNIEXPORT void JNICALL Java_hust_whno_SMLM_QC_1STORM_1Plug_lm_1FeedImageData(JNIEnv *env, jclass, jshortArray a, jint)
{
    PluginState &s = state();
    std::lock_guard<std::mutex> lk(s.mu);
    start_session(s, true);
    feed(s, env, a, 1);
}

JNIEXPORT void JNICALL Java_hust_whno_SMLM_QC_1STORM_1Plug_lm_1StopLocThread(JNIEnv *, jclass)
{
    PluginState &s = state();
    std::lock_guard<std::mutex> lk(s.mu);
    if (s.sess) report("qcs_session_finish", qcs_session_finish(s.sess));
}

JNIEXPORT void JNICALL Java_hust_whno_SMLM_QC_1STORM_1Plug_lm_1ReleaseLocResource(JNIEnv *, jclass)
{
    PluginState &s = state();
    std::lock_guard<std::mutex> lk(s.mu);
    release(s);
}

JNIEXPORT jint JNICALL Java_hust_whno_SMLM_QC_1STORM_1Plug_lm_1IsLocFinish(JNIEnv *, jclass)
{
    PluginState &s = state();
    std::lock_guard<std::mutex> lk(s.mu);
    return s.sess ? qcs_session_is_finished(s.sess) : 1;
}

JNIEXPORT jint JNICALL Java_hust_whno_SMLM_QC_1STORM_1Plug_lm_1GetMaxDispVal(JNIEnv *, jclass)
{
    PluginState &s = state();
    std::lock_guard<std::mutex> lk(s.mu);
    return max_disp(s);
}

JNIEXPORT jfloatArray JNICALL Java_hust_whno_SMLM_QC_1STORM_1Plug_lm_1GetSMLMImage(JNIEnv *env, jclass)
{
    PluginState &s = state();
    std::lock_guard<std::mutex> lk(s.mu);
    return sr_image(s, env);
}

JNIEXPORT jintArray JNICALL Java_hust_whno_SMLM_QC_1STORM_1Plug_lm_1GetSMLMImage3D(JNIEnv *env, jclass)
{
    PluginState &s = state();
    std::lock_guard<std::mutex> lk(s.mu);
    return sr_image_3d(s, env);
}

// synchronous extract + fit + statistics of a batch; the 8-float array of the reference
// (MM/QC_STORM_CPPDLL.cpp:857-924)
JNIEXPORT jfloatArray JNICALL Java_hust_whno_SMLM_QC_1STORM_1Plug_lm_1LocBatchedImg(JNIEnv *env, jclass, jshortArray img,
                                                                                    jint n)
{
    PluginState &s = state();
    std::lock_guard<std::mutex> lk(s.mu);
    para_init(s);
    float info[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    int rc = QCS_OK;
    if (!s.ctx) rc = qcs_ctx_create(&s.para, &s.ctx);
    else rc = qcs_ctx_set_params(s.ctx, &s.para);
    if (rc == QCS_OK && n > 0) {
        const long long cap = 2LL * QCS_MAX_POINT_NUM * (n > 0 ? n : 1);
        if ((long long)s.rows.size() < cap * QCS_REC_FLOATS) s.rows.resize((size_t)(cap * QCS_REC_FLOATS));
        jshort *px = (jshort *)env->GetPrimitiveArrayCritical(img, nullptr);
        if (px) {
            const long long rows = qcs_loc_batch_sync_info(s.ctx, (const uint16_t *)px, n, 1, s.rows.data(), cap, info, nullptr);
            env->ReleasePrimitiveArrayCritical(img, px, JNI_ABORT);
            if (rows < 0) {
                report("qcs_loc_batch_sync_info", (int)-rows);
                memset(info, 0, sizeof(info));
            }
        }
    } else {
        report("qcs_ctx_create", rc);
    }
    jfloatArray out = env->NewFloatArray(8);
    if (out) env->SetFloatArrayRegion(out, 0, 8, info);
    return out;
}

}  // extern "C"
