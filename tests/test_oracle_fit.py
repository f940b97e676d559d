"""Pin the oracle's BFGS fitters against mathematics: loss formula, noiseless recovery,
symmetries, FP64 agreement, closed-form CRLB and the Cramer-Rao bound."""
import ctypes as C

import numpy as np
import pytest

import oracle_lib as O
import qcstorm as Q


def _noiseless_roi(R, x0, y0, sigma, N, b, kadc=0.05, offset=100.0, sx=None, sy=None):
    """ROI record whose electron image is the sampled Gaussian (quantised to kadc)."""
    sx = sigma if sx is None else sx
    sy = sigma if sy is None else sy
    r = np.arange(R) + 0.5
    mu = N / (2 * np.pi * sx * sy) * np.exp(-((r[None, :] - x0) ** 2 / (2 * sx * sx) + (r[:, None] - y0) ** 2 / (2 * sy * sy))) + b
    rec = np.zeros((1, R * (R + 1)), np.uint16)
    rec[0, :R * R] = (np.round(mu / kadc) + offset).astype(np.uint16).ravel()
    rec[0, R * R:R * R + 2] = 100
    rec[0, R * R + 2] = 1
    return rec


def test_loss_matches_float64_formula():
    R = 7
    sp = Q.default_sim(width=R, height=R)
    rois, _ = Q.sim_rois_host(sp, 3, 20)
    params = Q.default_params(QE=1.0)
    w = O.wle(rois, R)
    for i in range(20):
        theta = np.array([8.0, 3.3, 3.7, 2.5, 1.0], np.float32)
        got = O.lib().orc_loss(C.byref(params), O.M_GS2D, O.p(rois[i:i + 1]), O.p(w[i:i + 1]), O.p(theta))
        q = np.maximum((rois[i, :49].astype(np.float64) - 100) * np.float64(np.float32(0.45)), 1.0).reshape(7, 7)
        c = np.arange(7) + 0.5
        s = 2.5 * 0.1
        I = 8.0 * 128 * np.exp(-s * ((c[None, :] - 3.3) ** 2 + (c[:, None] - 3.7) ** 2)) + 64.0
        mt = int(w[i, 5])
        ws = min(min(w[i, 1], w[i, 2]), min(w[i, 3], w[i, 4]), np.float32(7 / 2.0 / 2.35))
        if mt != 0:
            ws = min(ws / 1.2, 7 / 2.0 / 2.35)
        wgt = np.ones((7, 7)) if mt == 0 else np.exp(-((c[None, :] - 3.5) ** 2 + (c[:, None] - 3.5) ** 2) / (2 * ws * ws))
        ref = np.sum(wgt * (I - q - q * (np.log(I) - np.log(q))))
        assert abs(got - ref) <= 2e-5 * abs(ref) + 1e-3, (i, got, ref)


@pytest.mark.parametrize("x0,y0,sigma", [(3.5, 3.5, 1.3), (3.1, 3.9, 1.1), (4.2, 2.9, 1.5)])
def test_noiseless_recovery(x0, y0, sigma):
    R = 7
    rec = _noiseless_roi(R, x0, y0, sigma, N=5000.0, b=20.0)
    params = Q.default_params(QE=1.0, KAdc=0.05)
    w = O.wle(rec, R)
    th = O.fit_params(params, O.M_GS2D, rec, w, f64=True)[0]
    sg = np.sqrt(0.5 / (th[3] * 0.1))
    # the weighted, fixed-schedule fit lands close to the generating parameters
    assert abs(th[1] - x0) < 0.02 and abs(th[2] - y0) < 0.02 and abs(sg - sigma) < 0.05, (th, sg)
    amp = th[0] * 128
    assert abs(amp / (5000 / (2 * np.pi * sigma * sigma)) - 1) < 0.05


def test_mirror_and_transpose_symmetry_fp64():
    R = 7
    sp = Q.default_sim(width=R, height=R)
    rois, _ = Q.sim_rois_host(sp, 8, 400)
    params = Q.default_params(QE=1.0)
    P = R * R
    img = rois[:, :P].reshape(-1, R, R)
    mir = rois.copy()
    mir[:, :P] = img[:, :, ::-1].reshape(-1, P)
    tra = rois.copy()
    tra[:, :P] = img.transpose(0, 2, 1).reshape(-1, P)
    a = O.fit_params(params, O.M_GS2D, rois, O.wle(rois, R), f64=True)
    b = O.fit_params(params, O.M_GS2D, mir, O.wle(mir, R), f64=True)
    t = O.fit_params(params, O.M_GS2D, tra, O.wle(tra, R), f64=True)
    dm = np.abs((R - b[:, 1]) - a[:, 1])
    dt = np.maximum(np.abs(t[:, 1] - a[:, 2]), np.abs(t[:, 2] - a[:, 1]))
    print("mirror median %.2e p95 %.2e; transpose median %.2e" % (np.median(dm), np.percentile(dm, 95), np.median(dt)))
    assert np.median(dm) < 1e-6 and np.percentile(dm, 90) < 1e-3
    assert np.median(dt) < 1e-6 and np.percentile(dt, 90) < 1e-3


def test_fp32_schedule_tracks_fp64_schedule():
    R = 7
    sp = Q.default_sim(width=R, height=R)
    rois, _ = Q.sim_rois_host(sp, 9, 3000)
    params = Q.default_params(QE=1.0)
    w = O.wle(rois, R)
    a, _ = O.fit(params, rois, w)
    b, _ = O.fit_f64(params, rois, w)
    both = (a[:, 0] > 0) & (b[:, 0] > 0)
    d = np.abs(a[both, 1] - b[both, 1])
    print("FP32 vs FP64 |dx|: median %.2e p90 %.2e p99 %.2e" % (np.median(d), np.percentile(d, 90), np.percentile(d, 99)))
    # FP32 rounding noise of the fixed schedule: sub-1e-3 px typically, tails from bisection flips
    assert np.median(d) < 1e-3 and np.percentile(d, 90) < 2e-2
    assert np.mean((a[:, 0] > 0) == (b[:, 0] > 0)) > 0.99


def test_crlb_numeric_matches_closed_form():
    params = Q.default_params(QE=0.72, ReadNoise_e=1.3, PixelSize=100.0)
    # worked point of ML/CRLB in QC-STORM/Gaussian PSF 2D/CRLB_calc_main.m: N=2000, B=70, sigma=1.3 px
    loc = np.zeros((1, 12), np.float64)
    loc[0, 6] = 2000.0
    loc[0, 7] = 70.0
    loc[0, 4] = loc[0, 5] = 1.3
    out = loc.copy()
    O.lib().orc_crlb_f64(C.byref(params), O.p(out), 1)
    te, be = 2000.0 * 0.72, (70.0 * 0.72 * 0.72 + 1.3 ** 2) / 0.72
    g = np.arange(9) - 4.0
    X, Y = np.meshgrid(g, g)
    s = 1.3
    mu = te / (2 * np.pi * s * s) * np.exp(-(X ** 2 + Y ** 2) / (2 * s * s)) + be
    dmu = mu - be
    I = np.sum((dmu * X / (s * s)) ** 2 / mu)
    ref = np.sqrt(1 / I) * 100.0
    assert abs(out[0, 9] / ref - 1) < 2e-4 and abs(out[0, 10] / ref - 1) < 2e-4
    f32 = O.crlb(params, loc.astype(np.float32))
    assert abs(f32[0, 9] / ref - 1) < 1e-2


def test_rmse_against_fisher_bound():
    R = 7
    sp = Q.default_sim(width=R, height=R)
    rois, truth = Q.sim_rois_host(sp, 10, 3000)
    params = Q.default_params(QE=1.0)
    loc, _ = O.fit(params, rois, O.wle(rois, R))
    ok = loc[:, 0] > 0
    ex = loc[ok, 1] - (truth[ok, 0, 1] + 97)
    crlb = np.array([O.fisher_crlb(R, [truth[i, 0, 4], truth[i, 0, 1], truth[i, 0, 2], truth[i, 0, 6], truth[i, 0, 5]])[1]
                     for i in np.nonzero(ok)[0]])
    ratio = np.sqrt(np.mean(ex ** 2)) / np.sqrt(np.mean(crlb ** 2))
    print("reference schedule RMSE / CRLB =", ratio, "bias", ex.mean())
    assert 1.0 < ratio < 1.6 and abs(ex.mean()) < 0.01


def test_multi_emitter_cascade_resolves_pairs():
    R = 7
    sp = Q.default_sim(width=R, height=R, n_per_unit_max=2, min_sep=2.0, max_sep=2.5, photons_min=3000, photons_max=5000)
    rois, truth = Q.sim_rois_host(sp, 12, 600)
    rois[:, R * R + 4] = 1
    params = Q.default_params(QE=1.0, MultiEmitterFitEn=1)
    loc, st = O.fit(params, rois, O.wle(rois, R))
    pairs = truth[:, 1, 0] >= 0
    print(st, "pair ROIs", pairs.sum())
    # appended rows exist and every appended emitter lies near one of its ROI's true emitters
    assert st["n_rows"] > len(rois) and st["n_2e"] > 0
    app = loc[len(rois):]
    fr = app[:, 11].astype(int) - 1
    d = np.minimum(np.hypot(app[:, 1] - (truth[fr, 0, 1] + 97), app[:, 2] - (truth[fr, 0, 2] + 97)),
                   np.hypot(app[:, 1] - (truth[fr, 1, 1] + 97), app[:, 2] - (truth[fr, 1, 2] + 97)))
    assert np.median(d) < 0.3
