"""Pin the oracle's fixed-schedule fitters where a misreading of the reference would hide.

Every expected value below is fixed by the reference's own constants and rules (cited per test)
or by mathematics, not by re-running the oracle's formula:
  - the iteration schedule of each fitter through the number of loss evaluations it makes;
  - every acceptance edge of the one- and multi-emitter selections, on hand-built fit vectors
    placed on both sides of the edge (the expected accept/reject follows from the reference
    condition by construction);
  - the loss models of the elliptical one-emitter, three-emitter and elliptical two-emitter
    fitters through noiseless recovery (FP64, run to convergence and on the fixed schedule) and
    mirror / transpose symmetry;
  - the 15 x 15 CRLB of astigmatic rows against the closed-form Fisher information.

Reference files (DLL = MLEbfgs/bfgs_MLE_dll x64/bfgsMLE_dll/bfgsMLE_dll):
  DLL/bfgsMLE_Loc_LD/bfgsMLE_GS2D_core.h, bfgsMLE_AS3D_core.h,
  DLL/bfgsMLE_Loc_HD/bfgsMLE_GS2D_2Emitter_core.h, bfgsMLE_GS2D_3Emitter_core.h,
  bfgsMLE_AS3D_2Emitter_core.h, DLL/bfgsMLE_Loc_top/LocPrecCalc_GaussianCRLB.h
"""
import ctypes as C

import numpy as np
import pytest

import oracle_lib as O
import qcstorm as Q


def _roi(R, ems, b, kadc=0.05, offset=100.0, xc=100, yc=100, frame=1):
    """ROI record whose electron image is the sum of sampled Gaussians (x, y, N, sx, sy) + b."""
    r = np.arange(R) + 0.5
    mu = np.full((R, R), float(b))
    for (x0, y0, N, sx, sy) in ems:
        mu += N / (2 * np.pi * sx * sy) * np.exp(-((r[None, :] - x0) ** 2 / (2 * sx * sx) + (r[:, None] - y0) ** 2 / (2 * sy * sy)))
    rec = np.zeros((1, R * (R + 1)), np.uint16)
    rec[0, :R * R] = (np.round(mu / kadc) + offset).astype(np.uint16).ravel()
    rec[0, R * R] = xc
    rec[0, R * R + 1] = yc
    rec[0, R * R + 2] = frame
    return rec


def _header(R, xc=100, yc=100, frame=7):
    roi = np.zeros(R * (R + 1), np.uint16)
    roi[R * R:R * R + 3] = (xc, yc, frame)
    return roi


# ---------------------------------------------------------------------------------------------
# schedule: loss evaluations = 2P (first gradient) + It x (2 probes + 10 bisection steps)
#           + 2P (It - 1) (gradients after every iteration but the last); It and the 11-step
#           bisection: IterateNum_2D 8 (GS2D_core.h:25-26), IterateNum_AS3D 11 (AS3D_core.h:25-26),
#           IterateNum_2D_2E 13 (2Emitter_core.h:26-27), IterateNum_2D_3E 16, IterateNum_AS3D_2E 16
#           (AS3D_2Emitter_core.h:26-27); P = 5, 6, 8, 11, 11
SCHEDULE = [(O.M_GS2D, 7, 0, 0, 5, 8), (O.M_AS3D, 9, 1, 0, 6, 11), (O.M_2E, 7, 0, 1, 8, 13),
            (O.M_3E, 7, 0, 1, 11, 16), (O.M_AS3D2E, 9, 1, 1, 11, 16)]


@pytest.mark.parametrize("model,R,lt,multi,P,It", SCHEDULE)
def test_fixed_schedule_evaluations_and_iterations(model, R, lt, multi, P, It):
    params = Q.default_params(QE=1.0, KAdc=0.05, ROISize=R, LocType=lt, MultiEmitterFitEn=multi)
    rec = _roi(R, [(R / 2 + 0.3, R / 2 - 0.2, 5000, 1.3, 1.3)], 20)
    w = O.wle(rec, R, lt)
    O.loss_evals(reset=True)
    th = O.fit_params(params, model, rec, w)   # the model's own schedule
    n = O.loss_evals(reset=True)
    assert n == 2 * P + 12 * It + 2 * P * (It - 1), (n, P, It)
    # the default schedule is exactly It iterations: one less or one more gives other parameters
    same = O.fit_params(params, model, rec, w, iters=It)
    np.testing.assert_array_equal(th.view(np.uint32), same.view(np.uint32))
    for other in (It - 1, It + 1):
        assert not np.array_equal(O.fit_params(params, model, rec, w, iters=other), th)


# ---------------------------------------------------------------------------------------------
# one-emitter acceptance (bfgsMLE_GS2D_core.h:230-309): SNR > 4, 0.6 <= sigma <= R/1.5/2.35
# (R/2/2.35 in multi mode), x and y strictly inside (R/2 - 1.5, R/2 + 1.5); the fit vector holds
# A = PPho QE / 128, s = 5 / sigma^2 (sigma = sqrt(0.5 / (0.1 s))), B = Bkg QE / 64.
def _gs2d_x(pp, x, y, sigma, bg, qe=1.0):
    return np.array([pp * qe / 128, x, y, 5.0 / sigma ** 2, bg * qe / 64], np.float32)


@pytest.mark.parametrize("case,expect", [
    (dict(pp=20, bg=5), 0),          # SNR = 20 / sqrt(25) = 4 exactly: not > 4
    (dict(pp=21, bg=5), 1),
    (dict(sigma=1.98), 1), (dict(sigma=2.0), 0),          # R/1.5/2.35 = 1.986 at R = 7
    (dict(sigma=0.61), 1), (dict(sigma=0.59), 0),
    (dict(x=2.0), 0), (dict(x=2.01), 1), (dict(x=4.99), 1), (dict(x=5.0), 0),
    (dict(y=2.0), 0), (dict(y=4.99), 1),
    (dict(sigma=1.6, multi=1), 0), (dict(sigma=1.48, multi=1), 1),   # R/2/2.35 = 1.489 in multi mode
])
def test_gs2d_acceptance_edges(case, expect):
    c = dict(pp=500.0, x=3.5, y=3.5, sigma=1.3, bg=20.0, multi=0)
    c.update(case)
    params = Q.default_params(QE=1.0, ROISize=7, MultiEmitterFitEn=c["multi"])
    ok, row = O.select_1e(params, O.M_GS2D, _gs2d_x(c["pp"], c["x"], c["y"], c["sigma"], c["bg"]), _header(7, 300, 40, 9))
    assert ok == expect
    if ok:
        # row = [PPho, X, Y, 0, s, s, PPho 2 pi s^2, Bkg, SNR, 0, 0, frame] with X = x - R//2 + Xc
        s = c["sigma"]
        ref = [c["pp"], c["x"] - 3 + 300, c["y"] - 3 + 40, 0, s, s, c["pp"] * 2 * 3.141593 * s * s, c["bg"],
               c["pp"] / np.sqrt(c["pp"] + c["bg"]), 0, 0, 9]
        np.testing.assert_allclose(row, ref, rtol=2e-6, atol=1e-5)
    else:
        assert not row.any()


# elliptical one-emitter acceptance (bfgsMLE_AS3D_core.h:230-336): SNR (electrons) > 3.5,
# 0.8 < (sx + sy)/2 < R/1.5/2.35, x and y inside (R/2 - 2.5, R/2 + 2.5); z = sum p_k d^k with
# d = sx^2 - sy^2 (calibration p1 only here: z = p1 d)
def _as3d_x(pe, x, y, sx, sy, be):
    return np.array([pe / 128, x, y, 5.0 / sx ** 2, 5.0 / sy ** 2, be / 64], np.float32)


@pytest.mark.parametrize("case,expect", [
    (dict(pe=35, be=65), 0),        # 35 / sqrt(100) = 3.5 exactly
    (dict(pe=36, be=65), 1),
    (dict(sx=0.79, sy=0.79), 0), (dict(sx=0.81, sy=0.81), 1),
    (dict(sx=2.5, sy=2.62), 0), (dict(sx=2.5, sy=2.58), 1),      # R/1.5/2.35 = 2.553 at R = 9
    (dict(x=2.0), 0), (dict(x=2.01), 1), (dict(x=6.99), 1), (dict(x=7.0), 0),
])
def test_as3d_acceptance_edges(case, expect):
    c = dict(pe=2000.0, x=4.5, y=4.5, sx=1.2, sy=1.6, be=40.0)
    c.update(case)
    qe = 0.8
    params = Q.default_params(QE=qe, ROISize=9, LocType=1, p1_XGY=-111.11, p1_XLY=-111.11, p4_XGY=0.0, p4_XLY=0.0)
    ok, row = O.select_1e(params, O.M_AS3D, _as3d_x(c["pe"], c["x"], c["y"], c["sx"], c["sy"], c["be"]), _header(9, 50, 60, 3))
    assert ok == expect
    if ok:
        pp, bg = c["pe"] / qe, c["be"] / qe
        d = c["sx"] ** 2 - c["sy"] ** 2
        ref = [pp, c["x"] - 4 + 50, c["y"] - 4 + 60, -111.11 * d, c["sx"], c["sy"], pp * 2 * 3.141593 * c["sx"] * c["sy"], bg,
               c["pe"] / np.sqrt(c["pe"] + c["be"]), 0, 0, 3]
        np.testing.assert_allclose(row, ref, rtol=2e-5, atol=1e-4)


# ---------------------------------------------------------------------------------------------
# two-emitter 2D selection (bfgsMLE_GS2D_2Emitter_core.h:296-330): the first emitter is the one
# with the largest PPho x (R - distance to the centre); it is kept if SNR0 > 5.5 and
# 0.6 < sigma <= R/2/2.35; the second is added if distance > 1.5, SNR1 > 5.2, NN >= R/2,
# PPho1 > PPho0 / 2 and 1 < x1, y1 < R - 1.
def _me_x(ems, sigma, bg):
    x = []
    for (pp, ex, ey) in ems:
        x += [pp / 128, ex, ey]
    return np.array(x + [5.0 / sigma ** 2, bg / 64], np.float32)


BASE2 = dict(e0=(1000.0, 3.5, 3.5), e1=(600.0, 5.5, 3.5), sigma=1.2, bg=20.0, nn=5.0)


@pytest.mark.parametrize("case,expect", [
    ({}, 2),
    (dict(e1=(600.0, 5.0, 3.5)), 1), (dict(e1=(600.0, 5.01, 3.5)), 2),        # distance 1.5 exactly / 1.51
    (dict(e1=(500.0, 5.5, 3.5)), 1), (dict(e1=(501.0, 5.5, 3.5)), 2),        # PPho1 = PPho0 / 2 exactly
    (dict(e1=(400.0, 5.5, 3.5)), 1),                                          # /2 rule (a /3 rule would accept)
    (dict(nn=3.5), 2), (dict(nn=3.4999), 1),                                  # NN >= R/2 (3.5)
    (dict(e1=(600.0, 6.0, 3.5)), 1), (dict(e1=(600.0, 5.99, 3.5)), 2),        # x1 < R - 1
    (dict(e1=(600.0, 3.5, 1.0)), 1), (dict(e1=(600.0, 3.5, 1.01)), 2),        # y1 > 1
    (dict(e0=(100.0, 3.5, 3.5), e1=(65.0, 5.5, 3.5), bg=100.0), 1),           # SNR1 = 65/sqrt(165) = 5.06
    (dict(e0=(100.0, 3.5, 3.5), e1=(69.0, 5.5, 3.5), bg=100.0), 2),           # SNR1 = 69/13 = 5.31 (> 5.2, < 5.5)
    (dict(e0=(55.0, 3.5, 3.5), e1=(30.0, 5.5, 3.5), bg=45.0), 0),             # SNR0 = 55/10 = 5.5 exactly
    (dict(e0=(56.0, 3.5, 3.5), e1=(30.0, 5.5, 3.5), bg=45.0), 1),
    (dict(sigma=1.48), 2), (dict(sigma=1.5), 0), (dict(sigma=0.61), 2), (dict(sigma=0.59), 0),
])
def test_two_emitter_2d_selection_edges(case, expect):
    c = dict(BASE2)
    c.update(case)
    params = Q.default_params(QE=1.0, ROISize=7, MultiEmitterFitEn=1)
    k, row, row2 = O.select_me(params, O.M_2E, _me_x([c["e0"], c["e1"]], c["sigma"], c["bg"]), c["nn"], _header(7, 80, 90, 4))
    assert k == expect, (k, c)
    if k >= 1:
        pp, ex, ey = c["e0"]
        assert row[0] == np.float32(pp) and abs(row[1] - (ex - 3 + 80)) < 1e-5 and abs(row[2] - (ey - 3 + 90)) < 1e-5
        assert row[11] == 4
    if k == 2:
        pp, ex, ey = c["e1"]
        assert row2[0] == np.float32(pp) and abs(row2[1] - (ex - 3 + 80)) < 1e-5 and abs(row2[2] - (ey - 3 + 90)) < 1e-5
        assert abs(row2[8] - pp / np.sqrt(pp + c["bg"])) < 1e-4


def test_two_emitter_first_emitter_is_the_best_score():
    """The first emitter maximises PPho x (R - distance to the centre), not PPho."""
    params = Q.default_params(QE=1.0, ROISize=7, MultiEmitterFitEn=1)
    # scores 700 x 7 = 4900 (centre) vs 900 x (7 - 2.49) = 4059: the dimmer central emitter is first
    k, row, row2 = O.select_me(params, O.M_2E, _me_x([(700.0, 3.5, 3.5), (900.0, 5.99, 3.5)], 1.2, 20.0), 5.0,
                               _header(7, 0, 0, 1))
    assert k == 2 and row[0] == np.float32(700.0) and row2[0] == np.float32(900.0)
    # 2000 x 4.51 = 9020 > 4900: the off-centre emitter is first, and 700 < 2000 / 2 drops the second
    k, row, row2 = O.select_me(params, O.M_2E, _me_x([(700.0, 3.5, 3.5), (2000.0, 5.99, 3.5)], 1.2, 20.0), 5.0,
                               _header(7, 0, 0, 1))
    assert k == 1 and row[0] == np.float32(2000.0)


def _three_x(ems, sigma, bg):
    return _me_x(ems, sigma, bg)


# three-emitter 2D selection (bfgsMLE_GS2D_3Emitter_core.h:305-340): the same first-emitter rule;
# the second is the next best score; NN > R/2 strictly (WLE is off for this fitter, :54, so the
# fitter passes 1000)
@pytest.mark.parametrize("case,expect", [
    ({}, 2),
    (dict(nn=3.5), 1), (dict(nn=3.51), 2),
    (dict(e1=(400.0, 5.5, 3.5), e2=(500.0, 1.5, 3.5)), 1),   # second by score: 500 x 5 > 400 x 5, and 500 = 1000 / 2
    # second by score is the close one (450 x 6.5 = 2925 > 520 x 5 = 2600), rejected at distance 0.5;
    # picking the brighter 520 instead would accept it
    (dict(e1=(450.0, 4.0, 3.5), e2=(520.0, 1.5, 3.5)), 1),
])
def test_three_emitter_2d_selection_edges(case, expect):
    c = dict(e0=(1000.0, 3.5, 3.5), e1=(600.0, 5.5, 3.5), e2=(700.0, 1.5, 3.5), sigma=1.2, bg=20.0, nn=1000.0)
    c.update(case)
    params = Q.default_params(QE=1.0, ROISize=7, MultiEmitterFitEn=1)
    k, row, row2 = O.select_me(params, O.M_3E, _three_x([c["e0"], c["e1"], c["e2"]], c["sigma"], c["bg"]), c["nn"],
                               _header(7, 10, 10, 2))
    assert k == expect
    if k == 2:
        # scores: e1 600 x 5 = 3000, e2 700 x 5 = 3500 -> the second emitter is e2
        assert row2[0] == np.float32(700.0) and abs(row2[1] - (1.5 - 3 + 10)) < 1e-5


def test_three_emitter_fit_ignores_the_wle_neighbour_distance():
    R = 7
    rec = _roi(R, [(3.5, 3.5, 6000, 1.2, 1.2), (5.3, 3.4, 4000, 1.2, 1.2), (2.0, 5.3, 3000, 1.2, 1.2)], 20)
    params = Q.default_params(QE=1.0, KAdc=0.05, ROISize=R, MultiEmitterFitEn=1)
    w = O.wle(rec, R, 0)
    rows = []
    for nn in (0.5, 100.0):
        ww = w.copy()
        ww[:, 0] = nn
        ww[:, 5] = 2   # multi-emitter type: the cascade goes straight to the 2E fit
        th = O.fit_params(params, O.M_3E, rec, ww)
        rows.append(th)
    np.testing.assert_array_equal(rows[0].view(np.uint32), rows[1].view(np.uint32))
    # and the 2E fit does use it (weighted with NN)
    a = O.fit_params(params, O.M_2E, rec, w)
    ww = w.copy()
    ww[:, 0] = 0.5
    # the 2E weights do not depend on NN, only its selection does
    np.testing.assert_array_equal(O.fit_params(params, O.M_2E, rec, ww).view(np.uint32), a.view(np.uint32))


# elliptical two-emitter selection (bfgsMLE_AS3D_2Emitter_core.h:360-390): mean sigma in
# (0.8, R/1.9/2.35], SNR0 > 5.5, SNR1 > 5.2, distance > 1.5, NN > R/2, PPho1 > PPho0 / 3
def _as3d2_x(ems, bg):
    x = []
    for (pp, ex, ey, sx, sy) in ems:
        x += [pp / 128, ex, ey, 5.0 / sx ** 2, 5.0 / sy ** 2]
    return np.array(x + [bg / 64], np.float32)


@pytest.mark.parametrize("case,expect", [
    ({}, 2),
    (dict(e1=(400.0, 6.5, 4.5, 1.3, 1.3)), 2),      # PPho1 = 0.4 PPho0: /3 rule accepts (the 2D /2 rule would not)
    (dict(e1=(330.0, 6.5, 4.5, 1.3, 1.3)), 1),      # 330 < 1000 / 3
    (dict(e0=(1000.0, 4.5, 4.5, 2.0, 2.0)), 2),     # mean sigma 2.0 <= R/1.9/2.35 = 2.0157 at R = 9
    (dict(e0=(1000.0, 4.5, 4.5, 2.03, 2.03)), 0),
    (dict(e0=(1000.0, 4.5, 4.5, 0.79, 0.79)), 0), (dict(e0=(1000.0, 4.5, 4.5, 0.81, 0.81)), 2),
    (dict(e1=(600.0, 6.5, 4.5, 0.79, 0.79)), 1),
    (dict(nn=4.5), 1), (dict(nn=4.51), 2),          # NN > R/2 strictly
    (dict(e1=(600.0, 6.0, 4.5, 1.3, 1.3)), 1),      # distance 1.5 exactly
    (dict(e1=(600.0, 8.0, 4.5, 1.3, 1.3)), 1),      # x1 < R - 1 = 8
])
def test_two_emitter_elliptical_selection_edges(case, expect):
    c = dict(e0=(1000.0, 4.5, 4.5, 1.2, 1.6), e1=(600.0, 6.5, 4.5, 1.3, 1.3), bg=20.0, nn=1000.0)
    c.update(case)
    params = Q.default_params(QE=1.0, ROISize=9, LocType=1, MultiEmitterFitEn=1, p1_XGY=-111.11, p1_XLY=-111.11,
                              p4_XGY=0.0, p4_XLY=0.0)
    k, row, row2 = O.select_me(params, O.M_AS3D2E, _as3d2_x([c["e0"], c["e1"]], c["bg"]), c["nn"], _header(9, 20, 30, 5))
    assert k == expect, (k, c)
    if k >= 1:
        pp, ex, ey, sx, sy = c["e0"]
        assert abs(row[3] - (-111.11) * (sx * sx - sy * sy)) < 1e-3 and abs(row[4] - sx) < 1e-5 and abs(row[5] - sy) < 1e-5


# ---------------------------------------------------------------------------------------------
# loss models: noiseless ROIs.  Run to convergence (150 iterations of the same BFGS) the FP64
# fitters must reach the generating parameters; on the reference schedule they must come close.
def _sig(s):
    return np.sqrt(0.5 / (np.asarray(s) * 0.1))


@pytest.mark.parametrize("em", [(4.5, 4.5, 1.3, 1.7), (4.2, 4.9, 1.8, 1.1), (5.0, 4.1, 1.2, 1.25)])
def test_elliptical_one_emitter_noiseless_recovery(em):
    R = 9
    x0, y0, sx, sy = em
    rec = _roi(R, [(x0, y0, 8000, sx, sy)], 20)
    params = Q.default_params(QE=1.0, KAdc=0.05, ROISize=R, LocType=1)
    w = O.wle(rec, R, 1)
    for iters, tol in ((0, 3e-3), (150, 5e-4)):
        th = O.fit_params(params, O.M_AS3D, rec, w, iters=iters, f64=True)[0]
        assert abs(th[1] - x0) < tol and abs(th[2] - y0) < tol, (iters, th)
        assert abs(_sig(th[3]) - sx) < 5 * tol and abs(_sig(th[4]) - sy) < 5 * tol
        assert abs(th[5] * 64 - 20) < 0.1
        assert abs(th[0] * 128 / (8000 / (2 * np.pi * sx * sy)) - 1) < 5e-3


def _match(found, truth):
    """max over true emitters of the distance to the nearest found one"""
    return max(min(np.hypot(fx - tx, fy - ty) for fx, fy in found) for tx, ty in truth)


@pytest.mark.parametrize("ems", [
    [(3.5, 3.5, 6000, 1.2), (5.3, 3.4, 4000, 1.2), (2.0, 5.3, 3000, 1.2)],
    [(3.6, 3.3, 6000, 1.1), (1.9, 2.0, 4000, 1.1), (5.1, 5.2, 3500, 1.1)],
])
def test_three_emitter_noiseless_recovery(ems):
    R = 7
    rec = _roi(R, [(x, y, N, s, s) for (x, y, N, s) in ems], 20)
    params = Q.default_params(QE=1.0, KAdc=0.05, ROISize=R, MultiEmitterFitEn=1)
    w = O.wle(rec, R, 0)
    truth = [(x, y) for (x, y, _, _) in ems]
    for iters, tol in ((0, 0.05), (150, 1e-3)):
        th = O.fit_params(params, O.M_3E, rec, w, iters=iters, f64=True)[0]
        found = [(th[1], th[2]), (th[4], th[5]), (th[7], th[8])]
        assert _match(found, truth) < tol, (iters, th)
        assert abs(_sig(th[9]) - ems[0][3]) < 10 * tol and abs(th[10] * 64 - 20) < 0.5
    # converged photons: the amplitudes N / (2 pi s^2)
    amps = sorted(th[[0, 3, 6]] * 128)
    ref = sorted(N / (2 * np.pi * s * s) for (_, _, N, s) in ems)
    np.testing.assert_allclose(amps, ref, rtol=2e-3)


@pytest.mark.parametrize("ems", [
    [(4.5, 4.5, 6000, 1.2, 1.6), (6.5, 4.9, 5000, 1.5, 1.1)],
    [(4.3, 4.6, 6000, 1.4, 1.4), (2.6, 6.3, 4000, 1.1, 1.7)],
])
def test_elliptical_two_emitter_noiseless_recovery(ems):
    R = 9
    rec = _roi(R, ems, 20)
    params = Q.default_params(QE=1.0, KAdc=0.05, ROISize=R, LocType=1, MultiEmitterFitEn=1)
    w = O.wle(rec, R, 1)
    truth = [(x, y) for (x, y, _, _, _) in ems]
    for iters, tol in ((0, 0.25), (150, 1e-3)):
        th = O.fit_params(params, O.M_AS3D2E, rec, w, iters=iters, f64=True)[0]
        found = [(th[1], th[2]), (th[6], th[7])]
        assert _match(found, truth) < tol, (iters, th)
    # converged widths per emitter (matched by position)
    for (x, y, _, sx, sy) in ems:
        k = 0 if np.hypot(th[1] - x, th[2] - y) < np.hypot(th[6] - x, th[7] - y) else 1
        assert abs(_sig(th[5 * k + 3]) - sx) < 2e-3 and abs(_sig(th[5 * k + 4]) - sy) < 2e-3


def test_elliptical_fit_transpose_symmetry_fp64():
    """Transposing an astigmatic ROI swaps x <-> y and sx <-> sy of the FP64 fit."""
    R = 9
    sp = Q.default_sim(width=R, height=R, astig=1, photons_min=1000, photons_max=8000, bg_min=20, bg_max=80)
    rois, _ = Q.sim_rois_host(sp, 21, 300)
    P = R * R
    img = rois[:, :P].reshape(-1, R, R)
    tra = rois.copy()
    tra[:, :P] = img.transpose(0, 2, 1).reshape(-1, P)
    params = Q.default_params(QE=1.0, ROISize=R, LocType=1)
    a = O.fit_params(params, O.M_AS3D, rois, O.wle(rois, R, 1), f64=True)
    t = O.fit_params(params, O.M_AS3D, tra, O.wle(tra, R, 1), f64=True)
    d = np.maximum.reduce([np.abs(t[:, 1] - a[:, 2]), np.abs(t[:, 2] - a[:, 1]),
                           np.abs(_sig(t[:, 3]) - _sig(a[:, 4])), np.abs(_sig(t[:, 4]) - _sig(a[:, 3]))])
    print("transpose: median %.2e p90 %.2e" % (np.median(d), np.percentile(d, 90)))
    assert np.median(d) < 1e-6 and np.percentile(d, 90) < 1e-3


@pytest.mark.parametrize("model,R,lt", [(O.M_2E, 7, 0), (O.M_3E, 7, 0), (O.M_AS3D2E, 9, 1)])
def test_multi_emitter_fit_mirror_symmetry_fp64(model, R, lt):
    """A left-right mirrored ROI gives mirrored emitter positions, emitter by emitter (the
    starting blocks of the pre-estimation are placed symmetrically)."""
    sp = Q.default_sim(width=R, height=R, n_per_unit_max=3, min_sep=1.5, max_sep=3.0, astig=lt,
                       photons_min=2000, photons_max=6000)
    rois, _ = Q.sim_rois_host(sp, 31 + model, 200)
    P = R * R
    mir = rois.copy()
    mir[:, :P] = rois[:, :P].reshape(-1, R, R)[:, :, ::-1].reshape(-1, P)
    params = Q.default_params(QE=1.0, ROISize=R, LocType=lt, MultiEmitterFitEn=1)
    a = O.fit_params(params, model, rois, O.wle(rois, R, lt), f64=True)
    b = O.fit_params(params, model, mir, O.wle(mir, R, lt), f64=True)
    step = 5 if model == O.M_AS3D2E else 3
    K = 2 if model != O.M_3E else 3
    d = np.max([np.abs((R - b[:, step * k + 1]) - a[:, step * k + 1]) for k in range(K)], axis=0)
    print("mirror: median %.2e p75 %.2e" % (np.median(d), np.percentile(d, 75)))
    # ties among the starting blocks and bisection flips break the symmetry for a minority
    assert np.median(d) < 1e-6 and np.percentile(d, 60) < 1e-4


# ---------------------------------------------------------------------------------------------
# CRLB of astigmatic rows (LocPrecCalc_GaussianCRLB.h:64-243, 15 x 15 grid for AS3D): the
# second-difference Fisher information must equal the closed form sum (d mu / d x)^2 / mu on
# the 15 x 15 grid (and differ from the 9 x 9 one for wide PSFs).
@pytest.mark.parametrize("sx,sy", [(1.2, 2.4), (2.6, 1.1), (1.7, 1.7)])
def test_crlb_15x15_elliptical_closed_form(sx, sy):
    qe, rn, ps = 0.72, 1.3, 100.0
    params = Q.default_params(QE=qe, ReadNoise_e=rn, PixelSize=ps, LocType=1, ROISize=9)
    loc = np.zeros((1, 12), np.float64)
    loc[0, 6], loc[0, 7], loc[0, 4], loc[0, 5] = 3000.0, 50.0, sx, sy
    out = loc.copy()
    O.lib().orc_crlb_f64(C.byref(params), O.p(out), 1)
    te, be = 3000.0 * qe, (50.0 * qe * qe + rn * rn) / qe

    def closed(G):
        g = np.arange(G) - G // 2
        X, Y = np.meshgrid(g.astype(float), g.astype(float))
        mu0 = te / (2 * np.pi * sx * sy) * np.exp(-(X ** 2 / (2 * sx * sx) + Y ** 2 / (2 * sy * sy)))
        mu = mu0 + be
        Ix = np.sum((mu0 * X / (sx * sx)) ** 2 / mu)
        Iy = np.sum((mu0 * Y / (sy * sy)) ** 2 / mu)
        return np.sqrt(1 / Ix) * ps, np.sqrt(1 / Iy) * ps

    cx, cy = closed(15)
    assert abs(out[0, 9] / cx - 1) < 3e-4 and abs(out[0, 10] / cy - 1) < 3e-4, (out[0, 9:11], cx, cy)
    if max(sx, sy) > 2.0:
        nx, ny = closed(9)
        assert max(abs(out[0, 9] / nx - 1), abs(out[0, 10] / ny - 1)) > 1e-3
    f32 = O.crlb(params, loc.astype(np.float32))
    assert abs(f32[0, 9] / cx - 1) < 1e-2 and abs(f32[0, 10] / cy - 1) < 1e-2
