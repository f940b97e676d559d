"""ctypes binding of libqcstorm.so (the MI355X localization library).

The library's C-ABI is declared in include/qcstorm.h; this module mirrors its structs and
prototypes so Python callers (tests, benchmark, notebooks) can drive it with numpy arrays
(host API) or torch tensors (device API, HBM-resident data).  Loading fails loudly when the
built library is missing: there is no Python or CPU fallback for the localization path.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "lib", "libqcstorm.so")

QCS_OK = 0
QCS_ERR_ARG, QCS_ERR_HIP, QCS_ERR_CAPACITY, QCS_ERR_STATE, QCS_ERR_NOMEM, QCS_ERR_NODEVICE = 1, 2, 3, 4, 5, 6
LOCTYPE_GS2D, LOCTYPE_AS3D, LOCTYPE_DH3D = 0, 1, 2
MATH_EXACT, MATH_FAST = 0, 1
CFIT_INIT_HEUR, CFIT_INIT_GIVEN, CFIT_INIT_ADD = 0, 1, 2
CFIT_CONVERGED, CFIT_MAXITER, CFIT_STALLED, CFIT_FLAT, CFIT_PINNED = 1, 2, 4, 8, 16
# statistics of the multi-emitter paths (qcs_fit_warm_dev / qcs_fit_plain_stats_dev)
STAT_KEYS = ("final_k1", "final_k2", "final_k3", "fallbacks", "iters", "evals", "capped", "pred_k2", "pred_k3")
REC_FLOATS = 12
WLE_FLOATS = 6
TRUTH_FLOATS = 8
MAX_POINT_NUM = 20480 * 5
NOALLOC_MAX_ROIS = MAX_POINT_NUM
REC_FIELDS = ("PPho", "X", "Y", "Z", "SigX", "SigY", "TPho", "Bkg", "PSNR", "CrbX", "CrbY", "Frame")


def roi_words(R):
    return R * (R + 1)


class QcsParams(C.Structure):
    _fields_ = [
        ("Offset", C.c_float), ("KAdc", C.c_float), ("QE", C.c_float), ("ReadNoise_e", C.c_float),
        ("ROISize", C.c_int), ("LocType", C.c_int), ("MultiEmitterFitEn", C.c_int), ("WLEEn", C.c_int),
        ("ConsecFitEn", C.c_int), ("ConsecFit_DistanceTh_nm", C.c_float),
        ("PixelSize", C.c_float), ("PixelZoom", C.c_float), ("SNR_th", C.c_float), ("ColorMode_3D", C.c_int),
        ("ImageWidth", C.c_int), ("ImageHigh", C.c_int), ("TotalFrameNum", C.c_int),
        ("SRImageWidth", C.c_int), ("SRImageHigh", C.c_int),
        ("MinZDepth", C.c_float), ("MaxZDepth", C.c_float), ("ZDepthCorrFactor", C.c_float),
        ("p4_XGY", C.c_float), ("p3_XGY", C.c_float), ("p2_XGY", C.c_float), ("p1_XGY", C.c_float),
        ("p0_XGY", C.c_float),
        ("p4_XLY", C.c_float), ("p3_XLY", C.c_float), ("p2_XLY", C.c_float), ("p1_XLY", C.c_float),
        ("p0_XLY", C.c_float),
        ("RenderingMode", C.c_int), ("FixedLocPrec", C.c_float),
        ("FitMath", C.c_int), ("MaxROIPerBatch", C.c_int), ("BatchedImageNum", C.c_int), ("FlushROINum", C.c_int),
    ]


class QcsSimParams(C.Structure):
    _fields_ = [
        ("width", C.c_int), ("height", C.c_int), ("offset", C.c_float), ("kadc", C.c_float), ("qe", C.c_float),
        ("photons_min", C.c_float), ("photons_max", C.c_float), ("sigma_min", C.c_float), ("sigma_max", C.c_float),
        ("bg_min", C.c_float), ("bg_max", C.c_float), ("density_per_px", C.c_float), ("center_jitter", C.c_float),
        ("astig", C.c_int), ("z_min", C.c_float), ("z_max", C.c_float), ("s0", C.c_float), ("c_nm", C.c_float),
        ("d_nm", C.c_float), ("n_per_unit_max", C.c_int), ("min_sep", C.c_float), ("max_sep", C.c_float),
        ("density_max_per_px", C.c_float),
    ]


class QcsFitStats(C.Structure):
    _fields_ = [
        ("n_in", C.c_int), ("n_single", C.c_int), ("n_2e", C.c_int), ("n_3e", C.c_int), ("n_rejected", C.c_int),
        ("n_rows", C.c_int),
        ("FitRatio_1E", C.c_float), ("FitRatio_2E", C.c_float), ("FitRatio_3E", C.c_float),
        ("FitRatio_Final_1E", C.c_float), ("FitRatio_Final_2E", C.c_float), ("FitRatio_Final_3E", C.c_float),
        ("FitRatio_Final_4E", C.c_float),
    ]


class QcsTuning(C.Structure):
    _fields_ = [("max_iter", C.c_int * 3), ("tol", C.c_float), ("warm_pmin", C.c_float), ("queue", C.c_int),
                ("waves_per_cu", C.c_int), ("ftol", C.c_float), ("active_set", C.c_int), ("pin_stop", C.c_int)]


STAT_BINS = 1600


class QcsLocStats(C.Structure):
    _fields_ = [("hist_total_photon", C.c_int * STAT_BINS), ("hist_loc_prec", C.c_int * STAT_BINS),
                ("hist_peak_snr", C.c_int * STAT_BINS), ("hist_psf_width", C.c_int * STAT_BINS),
                ("valid", C.c_int), ("sum_loc_prec", C.c_double), ("sum_peak_snr", C.c_double),
                ("sum_psf_width", C.c_double), ("sum_background", C.c_double), ("mean_total_photon", C.c_float),
                ("mean_loc_prec", C.c_float), ("mean_peak_snr", C.c_float), ("mean_background", C.c_float),
                ("mean_psf_width", C.c_float), ("mean_psf_width_ctl", C.c_float), ("density_first_frame", C.c_int),
                ("density_last_frame", C.c_int), ("density_total", C.c_int), ("density_n0", C.c_int),
                ("density_n1", C.c_int), ("loc_density_2d", C.c_float)]


def loc_stats_dict(st):
    """qcs_loc_stats as a dict (histograms as numpy arrays)."""
    out = {}
    for k, _ in QcsLocStats._fields_:
        v = getattr(st, k)
        out[k] = np.frombuffer(v, dtype=np.int32).copy() if k.startswith("hist_") else v
    return out


class QcsAutotuneOpts(C.Structure):
    _fields_ = [("trials", C.c_int), ("min_select_agree", C.c_float), ("max_dx_p99", C.c_float)]


class QcsAutotuneEntry(C.Structure):
    _fields_ = [("tuning", QcsTuning), ("ms", C.c_float), ("select_agree", C.c_float), ("dx_p99", C.c_float),
                ("warm", C.c_int), ("within", C.c_int), ("chosen", C.c_int)]


class Graph:
    """A replayable HIP graph of one multi-emitter batch (qcs_graph_*)."""

    def __init__(self, h):
        self.h = h

    def launch(self, stream=None):
        check(lib().qcs_graph_launch(self.h, stream), "qcs_graph_launch")

    def close(self):
        if self.h:
            lib().qcs_graph_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


ONTIME_BINS = 10
RGB_BGR, RGB_RGB, RGB_BGRA, RGB_RGBA = 0, 1, 2, 3


class Consec:
    """Consecutive-frame merger (qcs_consec_*): rows of each call are held back until the next."""

    def __init__(self, ctx):
        h = C.c_void_p()
        check(lib().qcs_consec_create(ctx.h, C.byref(h)), "qcs_consec_create")
        self.h = h
        self.ctx = ctx   # the context must outlive the merger

    def fit_dev(self, d_rows, n, is_end, d_out, stream=None):
        """Returns (rows written to d_out, on-time ratios of the call's chains)."""
        m = C.c_int()
        ratio = np.zeros(ONTIME_BINS, np.float32)
        cap = d_out.shape[0] if d_out is not None else 0
        check(lib().qcs_consec_fit_dev(self.h, Context._dp(d_rows), n, int(is_end), Context._dp(d_out), cap,
                                       C.byref(m), C.c_void_p(ratio.ctypes.data), stream), "qcs_consec_fit_dev")
        return m.value, ratio

    def held(self):
        return lib().qcs_consec_held(self.h)

    def reset(self):
        check(lib().qcs_consec_reset(self.h), "qcs_consec_reset")

    def close(self):
        if self.h:
            lib().qcs_consec_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class QcsRoiSet(C.Structure):
    _fields_ = [("roi_size", C.c_int), ("n", C.c_int), ("d_rois", C.c_void_p), ("d_loc", C.c_void_p),
                ("d_sel", C.c_void_p), ("d_n", C.c_void_p)]


class QcsTiming(C.Structure):
    _fields_ = [
        ("extract_ms", C.c_double), ("fit_1e_ms", C.c_double), ("fit_me_ms", C.c_double), ("crlb_ms", C.c_double),
        ("compact_ms", C.c_double), ("render_ms", C.c_double),
        ("fit_1e_launches", C.c_longlong), ("fit_1e_rois", C.c_longlong), ("fit_me_launches", C.c_longlong),
        ("cfit_ms", C.c_double), ("nn_ms", C.c_double),
        ("cfit_launches", C.c_longlong), ("cfit_rois", C.c_longlong), ("nn_launches", C.c_longlong),
        ("drift_cc_ms", C.c_double), ("drift_cc_launches", C.c_longlong), ("drift_cc_pixels", C.c_longlong),
    ]


_P = C.POINTER
_vp, _i, _f, _u64, _ll = C.c_void_p, C.c_int, C.c_float, C.c_uint64, C.c_longlong

# name -> (restype, argtypes); every symbol declared in include/qcstorm.h
PROTOTYPES = {
    "qcs_params_default": (None, [_P(QcsParams)]),
    "qcs_params_update_sr_size": (None, [_P(QcsParams)]),
    "qcs_last_error": (C.c_char_p, []),
    "qcs_version": (C.c_char_p, []),
    "qcs_device_count": (_i, []),
    "qcs_test_fail_alloc_after": (None, [_i]),
    "qcs_ctx_create": (_i, [_P(QcsParams), _P(_vp)]),
    "qcs_ctx_set_params": (_i, [_vp, _P(QcsParams)]),
    "qcs_ctx_destroy": (_i, [_vp]),
    "qcs_extract_dev": (_i, [_vp, _vp, _i, _i, _vp, _vp, _i, _P(_i), _P(_i), _vp]),
    "qcs_wle_dev": (_i, [_vp, _vp, _i, _vp, _vp]),
    "qcs_fit_dev": (_i, [_vp, _vp, _vp, _i, _vp, _i, _P(_i), _vp]),
    "qcs_fit_dev_noalloc": (_i, [_vp, _vp, _vp, _i, _vp, _i, _vp, _vp]),
    "qcs_fit_stats_get": (_i, [_vp, _P(QcsFitStats)]),
    "qcs_crlb_dev": (_i, [_vp, _vp, _i, _vp]),
    "qcs_timing_enable": (_i, [_vp, _i]),
    "qcs_timing_get": (_i, [_vp, _P(QcsTiming)]),
    "qcs_compact_sorted_dev": (_i, [_vp, _vp, _i, _i, _i, _vp, _P(_i), _vp]),
    "qcs_compact_dev": (_i, [_vp, _vp, _i, _vp, _P(_i), _vp]),
    "qcs_render_dev": (_i, [_vp, _vp, _i, _vp, _vp, _vp]),
    "qcs_render_fresh_dev": (_i, [_vp, _vp, _i, _vp, _vp, _vp]),
    "qcs_fit_rois": (_i, [_vp, _vp, _vp, _i, _vp, _i, _P(_i)]),
    "qcs_cfit_dev": (_i, [_vp, _i, _vp, _i, _i, _vp, _vp, _vp, _vp, _vp]),
    "qcs_fit_mle_dev": (_i, [_vp, _vp, _i, _i, _vp, _vp, _vp]),
    "qcs_fit_mle_mixed_dev": (_i, [_vp, _P(QcsRoiSet), _i, _i, _vp]),
    "qcs_extract_mixed_dev": (_i, [_vp, _vp, _i, _i, _i, _P(QcsRoiSet), _vp, _vp]),
    "qcs_loc_mle_frames_dev": (_i, [_vp, _vp, _i, _i, _i, _i, _vp, _i, _vp, _vp]),
    "qcs_fit_plain_stats_dev": (_i, [_vp, _vp, _i, _i, _vp, _vp, _P(C.c_longlong), _vp]),
    "qcs_graph_create_mle": (_i, [_vp, _vp, _vp, _i, _i, _vp, _vp, _vp, _P(_vp)]),
    "qcs_drift_correct_dev": (_i, [_vp, _vp, _ll, _i, _vp, _vp, _i, _P(_i), _vp, _vp]),
    "qcs_consec_create": (_i, [_vp, _P(_vp)]),
    "qcs_consec_fit_dev": (_i, [_vp, _vp, _i, _i, _vp, _i, _P(_i), _vp, _vp]),
    "qcs_consec_held": (_i, [_vp]),
    "qcs_consec_reset": (_i, [_vp]),
    "qcs_consec_destroy": (_i, [_vp]),
    "qcs_graph_launch": (_i, [_vp, _vp]),
    "qcs_graph_create_frames": (_i, [_vp, _vp, _i, _vp, _i, _i, _vp, _i, _vp, _vp, _P(_vp)]),
    "qcs_loc_stats_dev": (_i, [_vp, _vp, _i, _P(QcsLocStats), _vp]),
    "qcs_autotune_mle_dev": (_i, [_vp, _vp, _vp, _i, _vp, _vp, _P(QcsAutotuneOpts), _P(QcsAutotuneEntry), _i, _P(_i), _vp]),
    "qcs_graph_destroy": (_i, [_vp]),
    "qcs_tuning_get": (_i, [_vp, _P(QcsTuning)]),
    "qcs_tuning_set": (_i, [_vp, _P(QcsTuning)]),
    "qcs_net_param_count": (_i, []),
    "qcs_net_create": (_i, [_i, _u64, _P(_vp)]),
    "qcs_net_destroy": (_i, [_vp]),
    "qcs_net_get_params": (_i, [_vp, _vp]),
    "qcs_net_set_params": (_i, [_vp, _vp]),
    "qcs_net_forward_dev": (_i, [_vp, _P(QcsParams), _vp, _i, _vp, _vp, _vp]),
    "qcs_net_forward_layers_dev": (_i, [_vp, _P(QcsParams), _vp, _i, _vp, _vp, _vp]),
    "qcs_net_grad_dev": (_i, [_vp, _P(QcsParams), _vp, _vp, _i, _vp, _P(_f), _vp]),
    "qcs_net_train_step_dev": (_i, [_vp, _P(QcsParams), _vp, _vp, _i, _f, _P(_f), _vp]),
    "qcs_label_dev": (_i, [_vp, _vp, _i, _vp, _vp, _vp]),
    "qcs_fit_warm_dev": (_i, [_vp, _vp, _vp, _i, _vp, _vp, _P(_ll), _vp]),
    "qcs_dev_gemm_nt_bf16": (_i, [_vp, _vp, _vp, _i, _i, _i, _vp]),
    "qcs_extract": (_i, [_vp, _vp, _i, _i, _vp, _vp, _i, _P(_i), _P(_i)]),
    "qcs_session_create": (_i, [_P(QcsParams), _P(_vp)]),
    "qcs_session_feed": (_i, [_vp, _vp, _i]),
    "qcs_session_finish": (_i, [_vp]),
    "qcs_session_is_finished": (_i, [_vp]),
    "qcs_session_fetch_locs": (_ll, [_vp, _vp, _ll]),
    "qcs_session_get_sr_image": (_i, [_vp, _vp]),
    "qcs_session_get_max_disp_val": (_i, [_vp, _P(_i)]),
    "qcs_session_destroy": (_i, [_vp]),
    "qcs_session_set_output": (_i, [_vp, C.c_char_p]),
    "qcs_session_get_sr_image3d": (_i, [_vp, _vp]),
    "qcs_render_rgb_dev": (_i, [_vp, _vp, _vp, _i, _f, _vp, _P(_f), _vp]),
    "qcs_session_get_stats": (_i, [_vp, _P(QcsLocStats), _vp]),
    "qcs_result_file_name": (_i, [_P(QcsParams), C.c_char_p, C.c_char_p, _i]),
    "qcs_loc_batch_sync": (_ll, [_vp, _vp, _i, _i, _vp, _ll]),
    "qcs_loc_batch_sync_info": (_ll, [_vp, _vp, _i, _i, _vp, _ll, _vp, _P(QcsLocStats)]),
    "qcs_sim_params_default": (None, [_P(QcsSimParams)]),
    "qcs_sim_rois_host": (_i, [_P(QcsSimParams), _u64, _i, _i, _vp, _vp]),
    "qcs_sim_rois_dev": (_i, [_P(QcsSimParams), _u64, _i, _i, _vp, _vp, _vp]),
    "qcs_sim_frame_emitters": (_i, [_P(QcsSimParams), _u64, _i, _vp, _i]),
    "qcs_sim_frames_host": (_i, [_P(QcsSimParams), _u64, _i, _i, _vp]),
    "qcs_sim_frames_dev": (_i, [_P(QcsSimParams), _u64, _i, _i, _vp, _vp]),
}

_lib = None


def lib():
    """Load libqcstorm.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libqcstorm.so not built at %s (run __graft_entry__.build())" % LIB_PATH)
        l = C.CDLL(LIB_PATH)
        for name, (res, args) in PROTOTYPES.items():
            fn = getattr(l, name, None)
            if fn is None:
                continue
            fn.restype = res
            fn.argtypes = args
        _lib = l
    return _lib


def missing_symbols():
    """Names declared in the header that the built library does not export."""
    l = C.CDLL(LIB_PATH)
    return [n for n in PROTOTYPES if getattr(l, n, None) is None]


class QcsError(RuntimeError):
    def __init__(self, rc, where):
        msg = lib().qcs_last_error().decode(errors="replace")
        super().__init__("%s failed (rc=%d): %s" % (where, rc, msg))
        self.rc = rc


def check(rc, where):
    if rc != QCS_OK:
        raise QcsError(rc, where)
    return rc


def default_params(**kw):
    p = QcsParams()
    lib().qcs_params_default(C.byref(p))
    for k, v in kw.items():
        if not hasattr(p, k):
            raise AttributeError("qcs_params has no field %r" % k)
        setattr(p, k, v)
    lib().qcs_params_update_sr_size(C.byref(p))
    return p


def result_file_name(params, image_name):
    """SetLocDataFileName: the result file name the reference derives from the image name."""
    n = lib().qcs_result_file_name(C.byref(params), image_name.encode(), None, 0)
    if n < 0:
        check(int(-n), "qcs_result_file_name")
    buf = C.create_string_buffer(n + 1)
    lib().qcs_result_file_name(C.byref(params), image_name.encode(), buf, n + 1)
    return buf.value.decode()


def default_sim(**kw):
    sp = QcsSimParams()
    lib().qcs_sim_params_default(C.byref(sp))
    for k, v in kw.items():
        if not hasattr(sp, k):
            raise AttributeError("qcs_sim_params has no field %r" % k)
        setattr(sp, k, v)
    return sp


def _ptr(a):
    return C.c_void_p(a.ctypes.data) if a is not None else None


def sim_rois_host(sp, seed, n, first_frame=1):
    R = sp.width
    M = max(1, sp.n_per_unit_max)
    rois = np.zeros((n, roi_words(R)), np.uint16)
    truth = np.zeros((n, M, TRUTH_FLOATS), np.float32)
    check(lib().qcs_sim_rois_host(C.byref(sp), seed, n, first_frame, _ptr(rois), _ptr(truth)), "qcs_sim_rois_host")
    return rois, truth


def sim_frames_host(sp, seed, first_frame, nframes):
    frames = np.zeros((nframes, sp.height, sp.width), np.uint16)
    check(lib().qcs_sim_frames_host(C.byref(sp), seed, first_frame, nframes, _ptr(frames)), "qcs_sim_frames_host")
    return frames


def frame_emitters(sp, seed, frame):
    n = lib().qcs_sim_frame_emitters(C.byref(sp), seed, frame, None, 0)
    t = np.zeros((max(n, 0), TRUTH_FLOATS), np.float32)
    lib().qcs_sim_frame_emitters(C.byref(sp), seed, frame, _ptr(t), n)
    return t


class Context:
    """A qcs_ctx bound to the current HIP device."""

    def __init__(self, params):
        self.params = params
        h = C.c_void_p()
        check(lib().qcs_ctx_create(C.byref(params), C.byref(h)), "qcs_ctx_create")
        self.h = h

    def close(self):
        if self.h:
            lib().qcs_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, params):
        check(lib().qcs_ctx_set_params(self.h, C.byref(params)), "qcs_ctx_set_params")
        self.params = params

    # ---- host-pointer API -------------------------------------------------------------------
    def fit_rois(self, rois, wle, loc_cap=None):
        n = rois.shape[0]
        cap = loc_cap or 2 * max(n, 1)
        loc = np.zeros((cap, REC_FLOATS), np.float32)
        rows = C.c_int()
        rois = np.ascontiguousarray(rois, np.uint16)
        wle = np.ascontiguousarray(wle, np.float32)
        check(lib().qcs_fit_rois(self.h, _ptr(rois), _ptr(wle), n, _ptr(loc), cap, C.byref(rows)), "qcs_fit_rois")
        return loc[:rows.value].copy()

    def extract(self, frames, start_frame=1, cap=None):
        frames = np.ascontiguousarray(frames, np.uint16)
        nf = frames.shape[0]
        R = self.params.ROISize
        cap = cap or MAX_POINT_NUM
        rois = np.zeros((cap, roi_words(R)), np.uint16)
        wle = np.zeros((cap, WLE_FLOATS), np.float32)
        n_out, n_found = C.c_int(), C.c_int()
        rc = lib().qcs_extract(self.h, _ptr(frames), nf, start_frame, _ptr(rois), _ptr(wle), cap, C.byref(n_out),
                               C.byref(n_found))
        if rc not in (QCS_OK, QCS_ERR_CAPACITY):
            check(rc, "qcs_extract")
        return rois[:n_out.value].copy(), wle[:n_out.value].copy(), n_found.value

    def timing_enable(self, on=True):
        check(lib().qcs_timing_enable(self.h, int(on)), "qcs_timing_enable")

    def timing(self):
        t = QcsTiming()
        check(lib().qcs_timing_get(self.h, C.byref(t)), "qcs_timing_get")
        return {k: getattr(t, k) for k, _ in QcsTiming._fields_}

    def stats(self):
        st = QcsFitStats()
        check(lib().qcs_fit_stats_get(self.h, C.byref(st)), "qcs_fit_stats_get")
        return {k: getattr(st, k) for k, _ in QcsFitStats._fields_}

    # ---- device-pointer API (torch tensors on cuda) -----------------------------------------
    @staticmethod
    def _dp(t):
        return C.c_void_p(t.data_ptr()) if t is not None else None

    def fit_dev(self, d_rois, d_wle, n, d_loc, stream=None):
        rows = C.c_int()
        cap = d_loc.shape[0]
        check(lib().qcs_fit_dev(self.h, self._dp(d_rois), self._dp(d_wle), n, self._dp(d_loc), cap, C.byref(rows),
                                stream), "qcs_fit_dev")
        return rows.value

    def fit_dev_noalloc(self, d_rois, d_wle, n, d_loc, d_n_rows, stream=None):
        """qcs_fit_dev_noalloc: the row count is left in the device int d_n_rows."""
        check(lib().qcs_fit_dev_noalloc(self.h, self._dp(d_rois), self._dp(d_wle), n, self._dp(d_loc), d_loc.shape[0],
                                        self._dp(d_n_rows), stream), "qcs_fit_dev_noalloc")

    def cfit_dev(self, K, d_rois, n, init_mode, d_init, d_theta, d_info=None, d_ll=None, stream=None):
        check(lib().qcs_cfit_dev(self.h, K, self._dp(d_rois), n, init_mode, self._dp(d_init), self._dp(d_theta),
                                 self._dp(d_info), self._dp(d_ll), stream), "qcs_cfit_dev")

    def fit_mle_dev(self, d_rois, n, kmax, d_loc, d_sel=None, stream=None):
        check(lib().qcs_fit_mle_dev(self.h, self._dp(d_rois), n, kmax, self._dp(d_loc), self._dp(d_sel), stream),
              "qcs_fit_mle_dev")

    def fit_mle_mixed_dev(self, sets, kmax, stream=None):
        """sets: list of (roi_size, d_rois, n, d_loc, d_sel or None[, d_n or None])."""
        arr = (QcsRoiSet * len(sets))()
        for i, st in enumerate(sets):
            R, d_rois, n, d_loc, d_sel = st[:5]
            d_n = st[5] if len(st) > 5 else None
            arr[i] = QcsRoiSet(R, n, d_rois.data_ptr(), d_loc.data_ptr(), d_sel.data_ptr() if d_sel is not None else None,
                               d_n.data_ptr() if d_n is not None else None)
        check(lib().qcs_fit_mle_mixed_dev(self.h, arr, len(sets), kmax, stream), "qcs_fit_mle_mixed_dev")

    def extract_mixed_dev(self, d_frames, nframes, start_frame, cap, set_caps, stream=None):
        """Mixed-size extraction into new device sets; returns ([(R, d_rois, count, d_n)] for R = 7, 9, 11)."""
        import torch
        dev = d_frames.device
        arr = (QcsRoiSet * 3)()
        outs = []
        for k, R in enumerate((7, 9, 11)):
            c = int(set_caps[k])
            d_rois = torch.zeros((max(c, 1), roi_words(R)), dtype=torch.int16, device=dev)
            d_n = torch.full((1,), -1, dtype=torch.int32, device=dev)
            arr[k] = QcsRoiSet(R, c, d_rois.data_ptr(), None, None, d_n.data_ptr())
            outs.append((R, d_rois, d_n))
        cnt = (C.c_int * 3)()
        check(lib().qcs_extract_mixed_dev(self.h, self._dp(d_frames), nframes, start_frame, cap, arr, cnt, stream),
              "qcs_extract_mixed_dev")
        return [(R, d_rois, cnt[k], d_n) for k, (R, d_rois, d_n) in enumerate(outs)]

    def loc_mle_frames_dev(self, d_frames, nframes, start_frame, cap, kmax, d_out, d_n_out, stream=None):
        check(lib().qcs_loc_mle_frames_dev(self.h, self._dp(d_frames), nframes, start_frame, cap, kmax, self._dp(d_out),
                                           d_out.shape[0], self._dp(d_n_out), stream), "qcs_loc_mle_frames_dev")

    def tuning(self):
        t = QcsTuning()
        check(lib().qcs_tuning_get(self.h, C.byref(t)), "qcs_tuning_get")
        return {"max_iter": list(t.max_iter), "tol": t.tol, "warm_pmin": t.warm_pmin, "queue": t.queue,
                "waves_per_cu": t.waves_per_cu, "ftol": t.ftol, "active_set": t.active_set, "pin_stop": t.pin_stop}

    def set_tuning(self, max_iter=None, tol=None, warm_pmin=None, queue=None, waves_per_cu=None, ftol=None,
                   active_set=None, pin_stop=None):
        t = QcsTuning()
        check(lib().qcs_tuning_get(self.h, C.byref(t)), "qcs_tuning_get")
        if max_iter is not None:
            for k in range(3):
                t.max_iter[k] = int(max_iter[k])
        if tol is not None:
            t.tol = tol
        if warm_pmin is not None:
            t.warm_pmin = warm_pmin
        if queue is not None:
            t.queue = int(queue)
        if waves_per_cu is not None:
            t.waves_per_cu = int(waves_per_cu)
        if ftol is not None:
            t.ftol = ftol
        if active_set is not None:
            t.active_set = int(active_set)
        if pin_stop is not None:
            t.pin_stop = int(pin_stop)
        check(lib().qcs_tuning_set(self.h, C.byref(t)), "qcs_tuning_set")

    def label_dev(self, d_rois, n, d_labels, d_sel=None, stream=None):
        check(lib().qcs_label_dev(self.h, self._dp(d_rois), n, self._dp(d_labels), self._dp(d_sel), stream), "qcs_label_dev")

    def fit_warm_dev(self, net, d_rois, n, d_loc, d_sel=None, stats=True, stream=None):
        st = (C.c_longlong * 9)()
        check(lib().qcs_fit_warm_dev(self.h, net.h, self._dp(d_rois), n, self._dp(d_loc), self._dp(d_sel),
                                     st if stats else None, stream), "qcs_fit_warm_dev")
        if not stats:
            return None
        return dict(zip(STAT_KEYS, list(st)))

    def autotune_dev(self, d_rois, n, net=None, trials=2, min_select_agree=0.99, max_dx_p99=0.02, stream=None):
        """Run the autotuner on n sample ROIs; returns the candidate table (list of dicts)."""
        import torch
        loc = torch.zeros((3 * n, 12), dtype=torch.float32, device=d_rois.device)
        sel = torch.zeros(n, dtype=torch.int32, device=d_rois.device)
        opts = QcsAutotuneOpts(trials, min_select_agree, max_dx_p99)
        cap = 64
        tab = (QcsAutotuneEntry * cap)()
        cnt = C.c_int()
        check(lib().qcs_autotune_mle_dev(self.h, net.h if net is not None else None, self._dp(d_rois), n, self._dp(loc),
                                         self._dp(sel), C.byref(opts), tab, cap, C.byref(cnt), stream), "qcs_autotune_mle_dev")
        out = []
        for e in tab[:min(cnt.value, cap)]:
            t = e.tuning
            out.append({"max_iter": list(t.max_iter), "ftol": t.ftol, "queue": t.queue, "waves_per_cu": t.waves_per_cu,
                        "warm_pmin": t.warm_pmin, "ms": e.ms, "select_agree": e.select_agree, "dx_p99": e.dx_p99,
                        "warm": bool(e.warm), "within": bool(e.within), "chosen": bool(e.chosen)})
        return out

    def loc_stats_dev(self, d_loc, n, stream=None):
        """Statistics of n localization rows on the device (waits for the result)."""
        st = QcsLocStats()
        check(lib().qcs_loc_stats_dev(self.h, self._dp(d_loc), n, C.byref(st), stream), "qcs_loc_stats_dev")
        return loc_stats_dict(st)

    def graph_frames(self, d_frames, nframes, d_start_frame, cap, kmax, d_out, d_n_out, stream=None):
        """A captured frame batch (qcs_loc_mle_frames_dev over fixed buffers, first frame number read
        from the device int d_start_frame at each launch); returns a Graph."""
        h = C.c_void_p()
        check(lib().qcs_graph_create_frames(self.h, self._dp(d_frames), nframes, self._dp(d_start_frame), cap, kmax,
                                            self._dp(d_out), d_out.shape[0], self._dp(d_n_out), stream, C.byref(h)),
              "qcs_graph_create_frames")
        return Graph(h)

    def graph_mle(self, d_rois, n, kmax, d_loc, d_sel=None, net=None, stream=None):
        """A captured multi-emitter batch (plain, or warm with `net`); returns a Graph."""
        h = C.c_void_p()
        check(lib().qcs_graph_create_mle(self.h, net.h if net is not None else None, self._dp(d_rois), n, kmax,
                                         self._dp(d_loc), self._dp(d_sel), stream, C.byref(h)), "qcs_graph_create_mle")
        return Graph(h)

    def fit_plain_stats_dev(self, d_rois, n, kmax, d_loc, d_sel=None, stream=None):
        """qcs_fit_mle_dev with statistics (waits for the result)."""
        st = (C.c_longlong * 9)()
        check(lib().qcs_fit_plain_stats_dev(self.h, self._dp(d_rois), n, kmax, self._dp(d_loc), self._dp(d_sel), st, stream),
              "qcs_fit_plain_stats_dev")
        return dict(zip(STAT_KEYS, list(st)))

    def wle_dev(self, d_rois, n, d_wle, stream=None):
        check(lib().qcs_wle_dev(self.h, self._dp(d_rois), n, self._dp(d_wle), stream), "qcs_wle_dev")

    def crlb_dev(self, d_loc, n, stream=None):
        check(lib().qcs_crlb_dev(self.h, self._dp(d_loc), n, stream), "qcs_crlb_dev")

    def extract_dev(self, d_frames, nframes, start_frame, d_rois, d_wle, cap, stream=None):
        n_out, n_found = C.c_int(), C.c_int()
        rc = lib().qcs_extract_dev(self.h, self._dp(d_frames), nframes, start_frame, self._dp(d_rois),
                                   self._dp(d_wle), cap, C.byref(n_out), C.byref(n_found), stream)
        if rc not in (QCS_OK, QCS_ERR_CAPACITY):
            check(rc, "qcs_extract_dev")
        return n_out.value, n_found.value

    def compact_sorted_dev(self, d_loc, n, first, end, d_out, stream=None):
        m = C.c_int()
        check(lib().qcs_compact_sorted_dev(self.h, self._dp(d_loc), n, first, end, self._dp(d_out), C.byref(m),
                                           stream), "qcs_compact_sorted_dev")
        return m.value

    def compact_dev(self, d_loc, n, d_out, stream=None):
        m = C.c_int()
        check(lib().qcs_compact_dev(self.h, self._dp(d_loc), n, self._dp(d_out), C.byref(m), stream),
              "qcs_compact_dev")
        return m.value

    def render_dev(self, d_loc, n, d_img, d_zimg=None, stream=None):
        check(lib().qcs_render_dev(self.h, self._dp(d_loc), n, self._dp(d_img), self._dp(d_zimg), stream),
              "qcs_render_dev")

    def render_fresh_dev(self, d_loc, n, d_img, d_zimg=None, stream=None):
        """Render into d_img (and d_zimg) replacing their contents (clear + render in one pass)."""
        check(lib().qcs_render_fresh_dev(self.h, self._dp(d_loc), n, self._dp(d_img), self._dp(d_zimg), stream),
              "qcs_render_fresh_dev")

    def drift_correct_dev(self, d_rows, n, corr_frames, shift_cap=None, stream=None):
        """Drift correction of n frame-sorted device rows in place.  Returns (frame count, per-frame x
        shifts, y shifts, group shifts as an (groups, 2) array)."""
        cap = shift_cap if shift_cap is not None else int(2 ** 20)
        fx = np.zeros(cap, np.float32)
        fy = np.zeros(cap, np.float32)
        gxy = np.full((512, 2), np.nan, np.float32)
        total = C.c_int()
        check(lib().qcs_drift_correct_dev(self.h, self._dp(d_rows), n, corr_frames, _ptr(fx), _ptr(fy), cap,
                                          C.byref(total), _ptr(gxy), stream), "qcs_drift_correct_dev")
        t = total.value
        g = int(np.count_nonzero(~np.isnan(gxy[:, 0])))
        return t, fx[:t].copy(), fy[:t].copy(), gxy[:g].copy()

    def render_rgb_dev(self, d_img, d_zimg, mode, d_out, bright_ratio=1.0, stream=None):
        """Colour-coded 3D image of the intensity / z planes into d_out; returns the brightness scale."""
        mx = C.c_float()
        check(lib().qcs_render_rgb_dev(self.h, self._dp(d_img), self._dp(d_zimg), mode, bright_ratio, self._dp(d_out),
                                       C.byref(mx), stream), "qcs_render_rgb_dev")
        return mx.value

    def loc_batch_sync(self, frames, start_frame=1, cap=None):
        """lm_LocBatchedImg: extract + fit + frame-sorted compaction of a group of frames (host arrays)."""
        frames = np.ascontiguousarray(frames, np.uint16)
        cap = cap or 2 * MAX_POINT_NUM
        out = np.zeros((cap, REC_FLOATS), np.float32)
        n = lib().qcs_loc_batch_sync(self.h, _ptr(frames), frames.shape[0], start_frame, _ptr(out), cap)
        if n < 0:
            check(int(-n), "qcs_loc_batch_sync")
        return out[:min(n, cap)].copy()

    def loc_batch_sync_info(self, frames, start_frame=1, cap=None):
        """lm_LocBatchedImg with its 8-float return array: (rows, info8, statistics dict)."""
        frames = np.ascontiguousarray(frames, np.uint16)
        cap = cap or 2 * MAX_POINT_NUM
        out = np.zeros((cap, REC_FLOATS), np.float32)
        info = np.zeros(8, np.float32)
        st = QcsLocStats()
        n = lib().qcs_loc_batch_sync_info(self.h, _ptr(frames), frames.shape[0], start_frame, _ptr(out), cap,
                                          _ptr(info), C.byref(st))
        if n < 0:
            check(int(-n), "qcs_loc_batch_sync_info")
        return out[:min(n, cap)].copy(), info, loc_stats_dict(st)


class Session:
    """Online localization session (the lm_* flow): feed frames, finish, fetch rows and the SR image."""

    def __init__(self, params):
        self.params = params
        h = C.c_void_p()
        check(lib().qcs_session_create(C.byref(params), C.byref(h)), "qcs_session_create")
        self.h = h
        # the library may fill SRImageWidth/SRImageHigh
        if params.SRImageWidth <= 16 or params.SRImageHigh <= 16:
            lib().qcs_params_update_sr_size(C.byref(self.params))

    def feed(self, frames):
        frames = np.ascontiguousarray(frames, np.uint16)
        check(lib().qcs_session_feed(self.h, _ptr(frames), frames.shape[0]), "qcs_session_feed")

    def finish(self):
        check(lib().qcs_session_finish(self.h), "qcs_session_finish")

    def is_finished(self):
        return bool(lib().qcs_session_is_finished(self.h))

    def wait(self, timeout_s=120.0, poll_s=0.005):
        import time
        t0 = time.time()
        while not self.is_finished():
            if time.time() - t0 > timeout_s:
                raise TimeoutError("session did not finish within %.0f s" % timeout_s)
            time.sleep(poll_s)

    def locs(self):
        n = lib().qcs_session_fetch_locs(self.h, None, 0)
        if n < 0:
            check(int(-n), "qcs_session_fetch_locs")
        out = np.zeros((n, REC_FLOATS), np.float32)
        m = lib().qcs_session_fetch_locs(self.h, _ptr(out), n)
        if m < 0:
            check(int(-m), "qcs_session_fetch_locs")
        return out

    def sr_image(self):
        img = np.zeros((self.params.SRImageHigh, self.params.SRImageWidth), np.float32)
        check(lib().qcs_session_get_sr_image(self.h, _ptr(img)), "qcs_session_get_sr_image")
        return img

    def max_disp_val(self):
        v = C.c_int()
        check(lib().qcs_session_get_max_disp_val(self.h, C.byref(v)), "qcs_session_get_max_disp_val")
        return v.value

    def sr_image3d(self):
        """lm_GetSMLMImage3D: the colour-coded 3D image, (SRImageHigh, SRImageWidth, 4) BGRA bytes."""
        img = np.zeros((self.params.SRImageHigh, self.params.SRImageWidth, 4), np.uint8)
        check(lib().qcs_session_get_sr_image3d(self.h, _ptr(img)), "qcs_session_get_sr_image3d")
        return img

    def set_output(self, path):
        """Append every flush's final rows to `path` (raw 12-float records); before the first feed."""
        check(lib().qcs_session_set_output(self.h, os.fsencode(path)), "qcs_session_set_output")

    def stats(self):
        """(statistics dict of the last flush's rows, on-time ratios of its chains)."""
        st = QcsLocStats()
        ratio = np.zeros(ONTIME_BINS, np.float32)
        check(lib().qcs_session_get_stats(self.h, C.byref(st), _ptr(ratio)), "qcs_session_get_stats")
        return loc_stats_dict(st), ratio

    def close(self):
        if self.h:
            lib().qcs_session_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Net:
    """The warm-start network (qcs_net): 7x7 ROIs, BF16 MFMA forward/backward, Adam."""

    def __init__(self, seed=0, roi_size=7):
        h = C.c_void_p()
        check(lib().qcs_net_create(roi_size, seed, C.byref(h)), "qcs_net_create")
        self.h = h
        self.nparam = lib().qcs_net_param_count()

    def close(self):
        if self.h:
            lib().qcs_net_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self):
        w = np.zeros(self.nparam, np.float32)
        check(lib().qcs_net_get_params(self.h, _ptr(w)), "qcs_net_get_params")
        return w

    def set_params(self, w):
        w = np.ascontiguousarray(w, np.float32)
        assert w.size == self.nparam
        check(lib().qcs_net_set_params(self.h, _ptr(w)), "qcs_net_set_params")

    def forward_dev(self, params, d_rois, n, d_out, d_norm=None, stream=None, layers=False):
        """Fused inference kernel, or (layers=True) the training path's layer-by-layer GEMMs."""
        fn = lib().qcs_net_forward_layers_dev if layers else lib().qcs_net_forward_dev
        check(fn(self.h, C.byref(params), C.c_void_p(d_rois.data_ptr()), n, C.c_void_p(d_out.data_ptr()),
                 C.c_void_p(d_norm.data_ptr()) if d_norm is not None else None, stream), "qcs_net_forward_dev")

    def grad_dev(self, params, d_rois, d_labels, n, stream=None):
        g = np.zeros(self.nparam, np.float32)
        loss = C.c_float()
        check(lib().qcs_net_grad_dev(self.h, C.byref(params), C.c_void_p(d_rois.data_ptr()),
                                     C.c_void_p(d_labels.data_ptr()), n, _ptr(g), C.byref(loss), stream),
              "qcs_net_grad_dev")
        return g, loss.value

    def train_step_dev(self, params, d_rois, d_labels, n, lr, stream=None, want_loss=True):
        loss = C.c_float()
        check(lib().qcs_net_train_step_dev(self.h, C.byref(params), C.c_void_p(d_rois.data_ptr()),
                                           C.c_void_p(d_labels.data_ptr()), n, lr,
                                           C.byref(loss) if want_loss else None, stream), "qcs_net_train_step_dev")
        return loss.value if want_loss else None
