"""ctypes binding of the CPU oracle (oracle/_build/liboracle.so).

Test infrastructure: only tests/, the smoke check and the benchmark's CPU-baseline leg use it,
always as the checker.  See oracle/orc.h for what each function restates.
"""
import ctypes as C
import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_SO = os.path.join(ORACLE_DIR, "_build", "liboracle.so")
sys.path.insert(0, os.path.join(ROOT, "qc-storm_amd"))

import qcstorm as Q  # noqa: E402

TMODE_DET, TMODE_LIBM = 0, 1
MODEL_SEPARABLE, MODEL_LITERAL = 0, 1
M_GS2D, M_AS3D, M_2E, M_3E, M_AS3D2E = 0, 1, 2, 3, 4
NPARAM = {M_GS2D: 5, M_AS3D: 6, M_2E: 8, M_3E: 11, M_AS3D2E: 11}


class OrcStats(C.Structure):
    _fields_ = [("n_in", C.c_int), ("n_single", C.c_int), ("n_2e", C.c_int), ("n_3e", C.c_int),
                ("n_rejected", C.c_int), ("n_rows", C.c_int)]


_vp, _i, _f, _P = C.c_void_p, C.c_int, C.c_float, C.POINTER
_PP = _P(Q.QcsParams)
PROTOS = {
    "orc_set_modes": (None, [_i, _i]),
    "orc_qexpf": (_f, [_f]),
    "orc_qlogf": (_f, [_f]),
    "orc_qexpf_v": (None, [_vp, _vp, C.c_long]),
    "orc_qlogf_v": (None, [_vp, _vp, C.c_long]),
    "orc_tlogf": (_f, [_f]),
    "orc_tlog_u": (_f, [_f, _f]),
    "orc_tlogf_v": (None, [_vp, _vp, C.c_long]),
    "orc_tlog_table": (None, [_vp]),
    "orc_tlogbf": (_f, [_f]),
    "orc_tlogb_u": (_f, [_f, _f]),
    "orc_tlogbf_v": (None, [_vp, _vp, C.c_long]),
    "orc_tlogb_table": (None, [_vp]),
    "orc_div_dh_fast": (_f, [_f]),
    "orc_div_dh_scan": (C.c_longlong, [C.c_uint32, C.c_uint32]),
    "orc_filter_taps": (None, [_i, _i, _i, _vp]),
    "orc_remove_background": (None, [_vp, _vp, _i, _i]),
    "orc_image_variance": (_f, [_vp, _i, _i]),
    "orc_conv_h": (None, [_vp, _vp, _i, _i, _vp]),
    "orc_conv_v": (None, [_vp, _vp, _i, _i, _vp]),
    "orc_find_rois": (_i, [_vp, _i, _i, _i, _f, _i, _i, _i, _vp, _vp, _i]),
    "orc_classify_rois": (None, [_vp, _vp, _i, _i, _i, _i, _vp]),
    "orc_cut_rois": (None, [_vp, _vp, _vp, _i, _i, _i, _i, _i, _i, _vp]),
    "orc_extract": (_i, [_PP, _vp, _i, _i, _vp, _vp, _i, _P(_i), _vp, _vp]),
    "orc_mixed_size": (_i, [_vp, _i, _i, _i, _i, _vp]),
    "orc_extract_mixed": (_i, [_PP, _vp, _i, _i, _i, _P(_vp), _P(_i), _P(_i), _vp, _P(_i), _P(_i)]),
    "orc_spline_interp": (None, [_vp, _vp, _i, _vp, _vp, _i]),
    "orc_wle_psf_width": (None, [_vp, _i, _i, _vp]),
    "orc_wle_nearest": (None, [_vp, _i, _i, _vp]),
    "orc_wle_classify": (None, [_vp, _i, _i, _i, _vp]),
    "orc_wle": (None, [_vp, _i, _i, _i, _vp]),
    "orc_fit": (_i, [_PP, _vp, _vp, _i, _vp, _i, _P(OrcStats)]),
    "orc_fit_f64": (_i, [_PP, _vp, _vp, _i, _vp, _i, _P(OrcStats)]),
    "orc_fit_omp": (_i, [_PP, _vp, _vp, _i, _vp, _i]),
    "orc_fit_params": (_i, [_PP, _i, _vp, _vp, _i, _vp, _i]),
    "orc_fit_params_f64": (_i, [_PP, _i, _vp, _vp, _i, _vp, _i]),
    "orc_loss": (_f, [_PP, _i, _vp, _vp, _vp]),
    "orc_loss_evals": (C.c_long, [_i]),
    "orc_select_1e": (_i, [_PP, _i, _vp, _vp, _vp]),
    "orc_select_me": (_i, [_PP, _i, _vp, _f, _vp, _vp, _vp]),
    "orc_crlb": (None, [_PP, _vp, _i]),
    "orc_crlb_f64": (None, [_PP, _vp, _i]),
    "orc_compact_sorted": (_i, [_vp, _i, _i, _i, _vp]),
    "orc_compact": (_i, [_vp, _i, _vp]),
    "orc_render": (None, [_PP, _vp, _i, _vp, _vp]),
    "orc_mle_f64": (_i, [_vp, _i, _i, _i, _vp, _i, _P(C.c_double)]),
    "orc_fisher_crlb_f64": (_i, [_i, _i, _i, _vp, _vp]),
    "orc_loc_stats": (_i, [_PP, _vp, _i, _P(Q.QcsLocStats)]),
    "orc_activation_density": (_f, [_f, _f, _f]),
    "orc_hist_mean": (_f, [_vp, _i, _f]),
    "orc_consec_links": (None, [_vp, _i, _f, _vp, _vp]),
    "orc_consec_create": (_vp, []),
    "orc_consec_destroy": (None, [_vp]),
    "orc_consec_held": (_i, [_vp]),
    "orc_consec_fit": (_i, [_PP, _vp, _vp, _i, _i, _vp, _i, _vp]),
    "orc_display_max_intensity": (_f, [_vp, C.c_longlong, _vp, _vp]),
    "orc_display_rgb": (None, [_PP, _vp, _vp, C.c_longlong, _f, _i, _vp]),
    "orc_drift_total_frame": (_i, [_vp, C.c_longlong]),
    "orc_drift_groups": (_i, [_i, _i, _vp, _vp]),
    "orc_drift_render": (None, [_PP, _vp, C.c_longlong, _i, _i, _vp]),
    "orc_drift_cc": (C.c_double, [_vp, _vp, _i, _i, _i, _i]),
    "orc_drift_peak": (None, [_vp, _i, _i, _P(_f), _P(_f)]),
    "orc_drift_interp": (None, [_i, _i, _vp, _vp, _vp, _vp, _vp, _vp]),
    "orc_drift_correct": (_i, [_PP, _vp, C.c_longlong, _i, _vp, _vp, _i, _P(_i), _vp, _vp]),
}

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            build()
        l = C.CDLL(ORACLE_SO)
        for name, (res, args) in PROTOS.items():
            fn = getattr(l, name)
            fn.restype = res
            fn.argtypes = args
        _lib = l
    return _lib


def p(a):
    return C.c_void_p(a.ctypes.data) if a is not None else None


def set_modes(tmode=TMODE_DET, model=MODEL_SEPARABLE):
    lib().orc_set_modes(tmode, model)


class modes:
    """Context manager selecting the oracle's transcendental and model-evaluation modes."""

    def __init__(self, tmode=TMODE_DET, model=MODEL_SEPARABLE):
        self.t, self.m = tmode, model

    def __enter__(self):
        set_modes(self.t, self.m)

    def __exit__(self, *a):
        set_modes(TMODE_DET, MODEL_SEPARABLE)


def wle(rois, R, loc_type=0):
    out = np.zeros((rois.shape[0], Q.WLE_FLOATS), np.float32)
    lib().orc_wle(p(np.ascontiguousarray(rois)), rois.shape[0], R, loc_type, p(out))
    return out


def fit(params, rois, wle_rec, cap=None):
    n = rois.shape[0]
    cap = cap or 2 * max(n, 1)
    loc = np.zeros((cap, Q.REC_FLOATS), np.float32)
    st = OrcStats()
    rows = lib().orc_fit(C.byref(params), p(np.ascontiguousarray(rois)), p(np.ascontiguousarray(wle_rec)), n,
                         p(loc), cap, C.byref(st))
    return loc[:rows].copy(), {k: getattr(st, k) for k, _ in OrcStats._fields_}


def fit_omp(params, rois, wle_rec, threads):
    n = rois.shape[0]
    loc = np.zeros((n, Q.REC_FLOATS), np.float32)
    lib().orc_fit_omp(C.byref(params), p(np.ascontiguousarray(rois)), p(np.ascontiguousarray(wle_rec)), n, p(loc),
                      threads)
    return loc


def fit_f64(params, rois, wle_rec, cap=None):
    n = rois.shape[0]
    cap = cap or 2 * max(n, 1)
    loc = np.zeros((cap, Q.REC_FLOATS), np.float64)
    st = OrcStats()
    rows = lib().orc_fit_f64(C.byref(params), p(np.ascontiguousarray(rois)), p(np.ascontiguousarray(wle_rec)), n,
                             p(loc), cap, C.byref(st))
    return loc[:rows].copy(), {k: getattr(st, k) for k, _ in OrcStats._fields_}


def fit_params(params, model, rois, wle_rec, iters=0, f64=False):
    n = rois.shape[0]
    P = NPARAM[model]
    th = np.zeros((n, P), np.float64 if f64 else np.float32)
    fn = lib().orc_fit_params_f64 if f64 else lib().orc_fit_params
    fn(C.byref(params), model, p(np.ascontiguousarray(rois)), p(np.ascontiguousarray(wle_rec)) if wle_rec is not None
       else None, n, p(th), iters)
    return th


def loss_evals(reset=True):
    return lib().orc_loss_evals(int(reset))


def select_1e(params, model, x, roi=None):
    """(accepted, row) of the one-emitter acceptance rule on the fit vector x."""
    R = params.ROISize
    roi = np.zeros(Q.roi_words(R), np.uint16) if roi is None else np.ascontiguousarray(roi, np.uint16)
    x = np.ascontiguousarray(x, np.float32)
    row = np.full(12, -9.0, np.float32)
    ok = lib().orc_select_1e(C.byref(params), model, p(x), p(roi), p(row))
    return ok, row


def select_me(params, model, x, nn=1000.0, roi=None):
    """(rows written 0/1/2, first row, second row) of the multi-emitter selection on x."""
    R = params.ROISize
    roi = np.zeros(Q.roi_words(R), np.uint16) if roi is None else np.ascontiguousarray(roi, np.uint16)
    x = np.ascontiguousarray(x, np.float32)
    row = np.full(12, -9.0, np.float32)
    row2 = np.full(12, -9.0, np.float32)
    k = lib().orc_select_me(C.byref(params), model, p(x), nn, p(roi), p(row), p(row2))
    return k, row, row2


def extract(params, frames, start_frame=1, cap=200000):
    frames = np.ascontiguousarray(frames, np.uint16)
    R = params.ROISize
    rois = np.zeros((cap, Q.roi_words(R)), np.uint16)
    w = np.zeros((cap, Q.WLE_FLOATS), np.float32)
    nb = (frames.shape[0] + max(params.BatchedImageNum, 1) - 1) // max(params.BatchedImageNum, 1)
    var = np.zeros(max(nb, 1), np.float32)
    found = C.c_int()
    n = lib().orc_extract(C.byref(params), p(frames), frames.shape[0], start_frame, p(rois), p(w), cap,
                          C.byref(found), None, p(var))
    return rois[:n].copy(), w[:n].copy(), found.value, var


def mixed_size(mark, h, x, y, wle7):
    """ROI size of the candidate at (x, y) of the stacked mark image (uint8, rows x w)."""
    mark = np.ascontiguousarray(mark, np.uint8)
    w7 = np.ascontiguousarray(wle7, np.float32)
    return lib().orc_mixed_size(p(mark), mark.shape[1], h, x, y, p(w7))


def extract_mixed(params, frames, start_frame=1, cand_cap=200000, caps=(200000, 200000, 200000)):
    """Mixed-size extraction: ([rois7, rois9, rois11], per-candidate sizes, candidates, found)."""
    frames = np.ascontiguousarray(frames, np.uint16)
    bufs = [np.zeros((max(caps[k], 1), Q.roi_words(7 + 2 * k)), np.uint16) for k in range(3)]
    ptrs = (C.c_void_p * 3)(*[b.ctypes.data for b in bufs])
    capa = (C.c_int * 3)(*caps)
    cnt = (C.c_int * 3)()
    sizes = np.zeros(max(cand_cap, 1), np.int32)
    nc, nf = C.c_int(), C.c_int()
    lib().orc_extract_mixed(C.byref(params), p(frames), frames.shape[0], start_frame, cand_cap, ptrs, capa, cnt, p(sizes),
                            C.byref(nc), C.byref(nf))
    return [bufs[k][:cnt[k]].copy() for k in range(3)], sizes[:nc.value].copy(), nc.value, nf.value


def crlb(params, loc):
    loc = np.ascontiguousarray(loc, np.float32).copy()
    lib().orc_crlb(C.byref(params), p(loc), loc.shape[0])
    return loc


def compact_sorted(loc, first, end):
    loc = np.ascontiguousarray(loc, np.float32)
    out = np.zeros_like(loc)
    m = lib().orc_compact_sorted(p(loc), loc.shape[0], first, end, p(out))
    return out[:m].copy()


def compact(loc):
    loc = np.ascontiguousarray(loc, np.float32)
    out = np.zeros_like(loc)
    m = lib().orc_compact(p(loc), loc.shape[0], p(out))
    return out[:m].copy()


def render(params, loc, with_z=False):
    img = np.zeros((params.SRImageHigh, params.SRImageWidth), np.float32)
    zimg = np.zeros_like(img) if with_z else None
    loc = np.ascontiguousarray(loc, np.float32)
    lib().orc_render(C.byref(params), p(loc), loc.shape[0], p(img), p(zimg))
    return img, zimg


def mle(img, K=1, astig=False, theta0=None, max_iter=200):
    img = np.ascontiguousarray(img, np.float64)
    R = img.shape[0]
    th = np.ascontiguousarray(theta0, np.float64).copy()
    ll = C.c_double()
    it = lib().orc_mle_f64(p(img), R, K, int(astig), p(th), max_iter, C.byref(ll))
    return th, it, ll.value


def fisher_crlb(R, theta, K=1, astig=False):
    th = np.ascontiguousarray(theta, np.float64)
    out = np.zeros_like(th)
    lib().orc_fisher_crlb_f64(R, K, int(astig), p(th), p(out))
    return out


def loc_stats(params, loc):
    loc = np.ascontiguousarray(loc, np.float32)
    st = Q.QcsLocStats()
    lib().orc_loc_stats(C.byref(params), p(loc), loc.shape[0], C.byref(st))
    return Q.loc_stats_dict(st)


def activation_density(r0, r1, radius_um):
    return lib().orc_activation_density(r0, r1, radius_um)


def hist_mean(h, pct):
    h = np.ascontiguousarray(h, np.int32)
    return lib().orc_hist_mean(p(h), h.shape[0], pct)


def consec_links(loc, th_px):
    loc = np.ascontiguousarray(loc, np.float32)
    n = loc.shape[0]
    back = np.zeros(max(n, 1), np.int32)
    fwd = np.zeros(max(n, 1), np.int32)
    lib().orc_consec_links(p(loc), n, th_px, p(back), p(fwd))
    return back[:n], fwd[:n]


class Consec:
    """The oracle's consecutive-frame merger (held-back group between calls)."""

    def __init__(self, params):
        self.params = params
        self.h = lib().orc_consec_create()

    def held(self):
        return lib().orc_consec_held(self.h)

    def fit(self, rows, is_end):
        rows = np.ascontiguousarray(rows, np.float32).reshape(-1, 12)
        cap = self.held() + rows.shape[0]
        out = np.zeros((max(cap, 1), 12), np.float32)
        ratio = np.zeros(10, np.float32)
        m = lib().orc_consec_fit(C.byref(self.params), self.h, p(rows), rows.shape[0], int(is_end), p(out), cap, p(ratio))
        assert m >= 0
        return out[:m].copy(), ratio

    def close(self):
        if self.h:
            lib().orc_consec_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


def drift_groups(total, corr):
    gs = np.zeros(512, np.int32)
    ge = np.zeros(512, np.int32)
    g = lib().orc_drift_groups(total, corr, p(gs), p(ge))
    return gs[:max(g, 0)].copy(), ge[:max(g, 0)].copy()


def drift_render(params, rows, srw, srh):
    rows = np.ascontiguousarray(rows, np.float32)
    img = np.zeros((srh, srw), np.float32)
    lib().orc_drift_render(C.byref(params), p(rows), rows.shape[0], srw, srh, p(img))
    return img


def drift_cc(img1, img2, sx, sy):
    img1 = np.ascontiguousarray(img1, np.float32)
    img2 = np.ascontiguousarray(img2, np.float32)
    return lib().orc_drift_cc(p(img1), p(img2), img1.shape[1], img1.shape[0], sx, sy)


def drift_peak(cc47, bias_x, bias_y):
    cc = np.ascontiguousarray(cc47, np.float64)
    sx, sy = C.c_float(), C.c_float()
    lib().orc_drift_peak(p(cc), bias_x, bias_y, C.byref(sx), C.byref(sy))
    return sx.value, sy.value


def drift_interp(total, gs, ge, gx, gy):
    gs, ge = np.ascontiguousarray(gs, np.int32), np.ascontiguousarray(ge, np.int32)
    gx, gy = np.ascontiguousarray(gx, np.float32), np.ascontiguousarray(gy, np.float32)
    fx = np.zeros(total, np.float32)
    fy = np.zeros(total, np.float32)
    lib().orc_drift_interp(total, gs.shape[0], p(gs), p(ge), p(gx), p(gy), p(fx), p(fy))
    return fx, fy


def drift_correct(params, rows, corr):
    """Returns (corrected rows, per-frame x shifts, y shifts, frame count, group x shifts, y shifts)."""
    rows = np.ascontiguousarray(rows, np.float32).copy()
    cap = int(rows[:, 11].max()) + 1 if rows.shape[0] else 1
    fx = np.zeros(cap, np.float32)
    fy = np.zeros(cap, np.float32)
    gx = np.zeros(512, np.float32)
    gy = np.zeros(512, np.float32)
    total = C.c_int()
    g = lib().orc_drift_correct(C.byref(params), p(rows), rows.shape[0], corr, p(fx), p(fy), cap, C.byref(total), p(gx), p(gy))
    assert g > 0, g
    t = total.value
    return rows, fx[:t].copy(), fy[:t].copy(), t, gx[:g].copy(), gy[:g].copy()


def display_max_intensity(img=None, hist=None):
    """(97% brightness level, 2048-bin histogram) of an SR image, or of given counts."""
    out = np.zeros(2048, np.int32)
    if hist is not None:
        h = np.ascontiguousarray(hist, np.int32)
        return lib().orc_display_max_intensity(None, 0, p(h), p(out)), out
    img = np.ascontiguousarray(img, np.float32)
    return lib().orc_display_max_intensity(p(img), img.size, None, p(out)), out


def display_rgb(params, img, zimg, scale, mode):
    img = np.ascontiguousarray(img, np.float32)
    zimg = np.ascontiguousarray(zimg, np.float32)
    nb = 3 if mode < 2 else 4
    out = np.zeros(img.size * nb, np.uint8)
    lib().orc_display_rgb(C.byref(params), p(img), p(zimg), img.size, scale, mode, p(out))
    return out.reshape(img.shape + (nb,))
