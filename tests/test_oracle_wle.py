"""Pin the oracle's WLE pre-estimation: spline, nearest neighbour, classification."""
import os

import numpy as np
import pytest

import oracle_lib as O
import qcstorm as Q

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def spline(xi, yi, xo):
    xi = np.ascontiguousarray(xi, np.float32)
    yi = np.ascontiguousarray(yi, np.float32)
    xo = np.ascontiguousarray(xo, np.float32)
    yo = np.zeros_like(xo)
    O.lib().orc_spline_interp(O.p(xi), O.p(yi), xi.size, O.p(xo), O.p(yo), xo.size)
    return yo


def test_spline_exact_on_linear_data_and_knots():
    xi = np.arange(7, dtype=np.float32)
    yi = 2.5 + 0.75 * xi
    xo = np.arange(61, dtype=np.float32) * np.float32(0.1)
    yo = spline(xi, yi, xo)
    assert np.allclose(yo, 2.5 + 0.75 * xo, atol=2e-6)
    y2 = np.array([0.1, 3.0, 7.5, 9.0, 6.0, 2.0, 0.4], np.float32)
    k = spline(xi, y2, xo)[::10]
    assert np.allclose(k, y2, atol=2e-6)


def test_spline_matches_scipy_natural_on_mat_fixture():
    from scipy.interpolate import CubicSpline
    d = np.load(os.path.join(GOLD, "spline_meandata.npz"))
    for key in ("MeanData1", "MeanData2"):
        y = d[key].astype(np.float32).ravel()
        xi = np.arange(y.size, dtype=np.float32)
        xo = np.arange((y.size - 1) * 10 + 1, dtype=np.float32) * np.float32(0.1)
        ref = CubicSpline(xi.astype(np.float64), y.astype(np.float64), bc_type="natural")(xo.astype(np.float64))
        got = spline(xi, y, xo)
        assert np.max(np.abs(got - ref)) <= 1e-5 * max(1.0, np.abs(ref).max())


def _psf_roi(R, sigma, x0, y0, amp=2000.0, bg=10.0):
    r = np.arange(R) + 0.5
    img = amp * np.exp(-((r[None, :] - x0) ** 2 + (r[:, None] - y0) ** 2) / (2 * sigma ** 2)) + bg
    rec = np.zeros((1, R * (R + 1)), np.uint16)
    rec[0, :R * R] = np.round(img).astype(np.uint16).ravel()
    return rec


@pytest.mark.parametrize("sigma", [1.0, 1.3, 1.6])
def test_psf_width_recovers_sigma_of_centered_gaussian(sigma):
    R = 7
    w = np.zeros((1, 6), np.float32)
    O.lib().orc_wle_psf_width(O.p(_psf_roi(R, sigma, 3.5, 3.5)), 1, R, O.p(w))
    # marginal of a centred Gaussian: both one-sided widths ~ sigma (truncated at the ROI edge)
    assert np.allclose(w[0, 1:5], sigma, rtol=0.15), w


def test_psf_width_mirror_symmetry():
    R = 9
    rec = _psf_roi(R, 1.4, 4.2, 4.9)
    mir = rec.copy()
    mir[0, :R * R] = rec[0, :R * R].reshape(R, R)[:, ::-1].ravel()
    a = np.zeros((1, 6), np.float32)
    b = np.zeros((1, 6), np.float32)
    O.lib().orc_wle_psf_width(O.p(rec), 1, R, O.p(a))
    O.lib().orc_wle_psf_width(O.p(mir), 1, R, O.p(b))
    assert np.allclose([a[0, 1], a[0, 2]], [b[0, 2], b[0, 1]], rtol=1e-3)
    assert np.allclose(a[0, 3:5], b[0, 3:5], rtol=1e-6)


def test_nearest_neighbour_brute_force():
    rng = np.random.default_rng(3)
    R, n = 7, 500
    rois = np.zeros((n, R * (R + 1)), np.uint16)
    f = rng.integers(1, 6, n)
    x = rng.integers(0, 400, n)
    y = rng.integers(0, 400, n)
    P = R * R
    rois[:, P], rois[:, P + 1], rois[:, P + 2], rois[:, P + 3] = x, y, f % 65536, f // 65536
    w = np.zeros((n, 6), np.float32)
    O.lib().orc_wle_nearest(O.p(rois), n, R, O.p(w))
    d2 = (x[:, None] - x[None]) ** 2 + (y[:, None] - y[None]) ** 2
    same = (f[:, None] == f[None]) & ~np.eye(n, dtype=bool)
    best = np.where(same, d2, 10 ** 9).min(1).astype(np.float64)
    best = np.where(best < 1e8, best, 1e8)
    assert np.array_equal(w[:, 0], np.sqrt(best.astype(np.float32)))


def test_classification_rule():
    R = 7
    th = np.float32(np.float32(7 / np.float32(2.1)) / np.float32(2.35))
    cases = [  # nn, sL, sR, sU, sD, det -> type
        (50.0, 1.0, 1.0, 1.0, 1.0, 0, 0),      # isolated, clean, narrow -> MLE
        (50.0, 1.0, 1.5, 1.0, 1.0, 0, 1),      # contaminated left/right -> WLE
        (6.0, 1.0, 1.0, 1.0, 1.0, 0, 1),       # neighbour inside ROI but > R/2+1 -> WLE
        (4.0, 1.0, 1.0, 1.0, 1.0, 0, 2),       # neighbour closer than R/2+1 -> multi
        (50.0, 1.0, 1.0, 1.0, 1.0, 1, 2),      # detector says multi -> multi
        (50.0, 2.0, 2.0, 2.0, 2.0, 0, 2),      # too wide -> multi
        (50.0, th, th, th, th, 0, 0),          # on the width threshold -> MLE
    ]
    rois = np.zeros((len(cases), R * (R + 1)), np.uint16)
    w = np.zeros((len(cases), 6), np.float32)
    for i, c in enumerate(cases):
        w[i, :5] = c[:5]
        rois[i, R * R + 4] = c[5]
    O.lib().orc_wle_classify(O.p(rois), len(cases), R, 0, O.p(w))
    assert list(w[:, 5].astype(int)) == [c[6] for c in cases]
