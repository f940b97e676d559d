"""Pinning the oracle's drift correction (oracle/orc_drift.c): the group rule and the correlation
against hand cases and numpy, the peak refinement against a Gaussian peak of known centre, the
interpolation against its closed form, the render against the sampled Gaussian, and the whole
correction against a planted linear drift."""
import numpy as np
import pytest

import oracle_lib as O
import qcstorm as Q
from drift_data import drifting_rows


def test_groups():
    gs, ge = O.drift_groups(100, 20)
    assert gs.tolist() == [1, 11, 21, 31, 41, 51, 61, 71, 81] and ge.tolist() == [20, 30, 40, 50, 60, 70, 80, 90, 100]
    gs, ge = O.drift_groups(105, 20)       # remainder below half a group: joins the last group
    assert ge[-1] == 105 and len(gs) == 9
    gs, ge = O.drift_groups(115, 20)       # remainder of half a group or more: its own group
    assert (gs[-1], ge[-1]) == (96, 115) and len(gs) == 10
    assert O.drift_groups(15, 20)[0].size == 0   # fewer frames than a group


def test_cc_matches_numpy():
    rng = np.random.default_rng(0)
    a = rng.random((40, 56)).astype(np.float32)
    b = rng.random((40, 56)).astype(np.float32)
    for sx, sy in [(0, 0), (3, -2), (-7, 5), (30, 0), (0, -39), (56, 0)]:
        h, w = a.shape
        ys, xs = np.mgrid[0:h, 0:w]
        ok = (xs + sx >= 0) & (xs + sx < w) & (ys + sy >= 0) & (ys + sy < h)
        want = float((a.astype(np.float64)[ok] * b.astype(np.float64)[(ys + sy)[ok], (xs + sx)[ok]]).sum())
        assert O.drift_cc(a, b, sx, sy) == pytest.approx(want, rel=1e-5, abs=1e-9)


@pytest.mark.parametrize("cx,cy", [(0.3, -0.2), (-0.45, 0.1), (5.2, -7.7)])
def test_peak_centre_of_mass(cx, cy):
    y, x = np.mgrid[0:47, 0:47]
    cc = 1e6 * np.exp(-((x - 23 - cx) ** 2 + (y - 23 - cy) ** 2) / (2 * 2.0 ** 2))
    sx, sy = O.drift_peak(cc, 10, -4)
    assert sx == pytest.approx(10 + cx, abs=2e-3) and sy == pytest.approx(-4 + cy, abs=2e-3)


def test_interpolation_closed_form():
    gs, ge = O.drift_groups(100, 20)
    gx = np.array([0.0, 1.0, 2.0, 3.0, 4.0, 5.0, 6.0, 7.0, 8.0], np.float32)
    fx, fy = O.drift_interp(100, gs, ge, gx, -gx)
    centres = (gs + ge) // 2                       # 10, 20, ..., 90
    assert fx[0] == 0.0
    np.testing.assert_allclose(fx[centres], gx, atol=1e-6)
    # between centres and before the first: linear; group 0's centre carries its (zero) shift,
    # frames before it go from 0 at index 0
    np.testing.assert_allclose(fx[15], 0.5, atol=1e-6)
    np.testing.assert_allclose(fx[99], 8.0 + 9 * 0.1, atol=1e-5)   # last slope continued
    np.testing.assert_allclose(fy, -fx)


def test_render_sampled_gaussian():
    p = Q.default_params(PixelSize=100.0, SNR_th=5.0)
    r = np.zeros((1, 12), np.float32)
    r[0, [0, 1, 2, 8, 9, 11]] = [500, 10.3, 7.6, 20, 10.0, 1]
    img = O.drift_render(p, r, 160, 160)
    sig = 10.0 / (100.0 / 8)
    rx, ry = np.float32(10.3) * 8, np.float32(7.6) * 8
    y, x = np.mgrid[0:160, 0:160]
    want = np.exp(-((x - rx) ** 2 + (y - ry) ** 2) / (2 * sig * sig))
    win = (np.abs(x - int(rx)) <= 3) & (np.abs(y - int(ry)) <= 3) & (x >= int(rx) - 3) & (y >= int(ry) - 3)
    np.testing.assert_allclose(img[win], want[win], rtol=2e-6, atol=1e-12)
    assert not img[~win].any()
    for k, v in [(8, 4.0), (0, 1.0), (1, 0.6)]:   # PSNR below SNR_th, PPho <= 1, too close to the edge
        q = r.copy()
        q[0, k] = v
        assert not O.drift_render(p, q, 160, 160).any()
    q = r.copy()
    q[0, 9] = 80.0                                 # CRLB outside (0.01, 35): rendered 35 nm wide
    assert O.drift_render(p, q, 160, 160).max() == pytest.approx(1.0, abs=0.02)


def test_correction_recovers_planted_drift():
    W = H = 48
    rows, dx, dy, _ = drifting_rows(W, H, 60, 70, 1, wobble=0.0)
    p = Q.default_params(ImageWidth=W, ImageHigh=H, PixelSize=100.0, SNR_th=5.0)
    out, fx, fy, total, gx, gy = O.drift_correct(p, rows, 20)
    assert total == 60 and gx.size == 5
    gs, ge = O.drift_groups(60, 20)
    mean_dx = np.array([dx[s - 1:e].mean() for s, e in zip(gs, ge)])
    mean_dy = np.array([dy[s - 1:e].mean() for s, e in zip(gs, ge)])
    # the centre of mass over a 12 SR-pixel radius includes correlation with neighbouring
    # structure: the method's accuracy here is a few hundredths of a pixel
    np.testing.assert_allclose(gx, mean_dx - mean_dx[0], atol=0.04)
    np.testing.assert_allclose(gy, mean_dy - mean_dy[0], atol=0.04)
    assert np.polyfit(np.arange(5), gx, 1)[0] == pytest.approx(0.1, abs=0.01)     # 10 frames x 0.01 px
    assert np.polyfit(np.arange(5), gy, 1)[0] == pytest.approx(-0.06, abs=0.01)
    # rows are shifted by their frame's interpolated shift; other fields untouched
    f = rows[:, 11].astype(int)
    np.testing.assert_array_equal(out[:, 1], rows[:, 1] - fx[f - 1])
    np.testing.assert_array_equal(out[:, 2], rows[:, 2] - fy[f - 1])
    np.testing.assert_array_equal(out[:, [0, 3, 4, 5, 6, 7, 8, 9, 10, 11]], rows[:, [0, 3, 4, 5, 6, 7, 8, 9, 10, 11]])
