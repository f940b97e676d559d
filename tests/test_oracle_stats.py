"""Pinning the oracle's localization statistics (oracle/orc_stats.c): histograms and sums against
an independent numpy restatement, neighbour counts against a k-d tree and lattices with known
answers, the frame-range rule, and the density and histogram-mean formulas against closed forms."""
import numpy as np
import pytest
from scipy.spatial import cKDTree

import oracle_lib as O
import qcstorm as Q


def _rows(n, seed, nframes=10, W=256, H=256):
    rng = np.random.default_rng(seed)
    loc = np.zeros((n, 12), np.float32)
    loc[:, 0] = rng.uniform(0, 500, n)
    loc[:, 1] = rng.uniform(0, W, n)
    loc[:, 2] = rng.uniform(0, H, n)
    loc[:, 4] = rng.uniform(0.0, 2.5, n)
    loc[:, 5] = rng.uniform(0.0, 2.5, n)
    loc[:, 6] = rng.uniform(0, 12000, n)
    loc[:, 7] = rng.uniform(0, 60, n)
    loc[:, 8] = rng.uniform(0, 60, n)
    loc[:, 9] = rng.choice([np.nan, -1.0, 0.005, 3.0, 12.0, 30.0, 45.0], n) * rng.uniform(0.5, 1.5, n)
    loc[:, 10] = rng.uniform(0, 40, n)
    loc[:, 11] = np.sort(rng.integers(1, nframes + 1, n))
    return loc


def _numpy_hist(loc):
    f = np.float32
    tp, bg, snr = loc[:, 6], loc[:, 7], loc[:, 8]
    sg = (loc[:, 4] + loc[:, 5]) / f(2)
    lp = np.sqrt(loc[:, 9] * loc[:, 9] + loc[:, 10] * loc[:, 10])
    with np.errstate(invalid="ignore"):
        ok = (tp > 30) & (bg > 2) & (snr >= 3) & (sg >= f(0.01)) & (lp > f(0.01)) & (lp < 55)
    def h(v, gap):
        b = np.clip((v[ok] / f(gap)).astype(np.int64), 0, 1599)
        return np.bincount(b, minlength=1600)
    hists = {"hist_total_photon": h(tp, f(10000) / 400), "hist_loc_prec": h(lp, f(60) / 400),
             "hist_peak_snr": h(snr, f(50) / 400), "hist_psf_width": h(sg, f(5) / 400)}
    sums = {"sum_loc_prec": lp[ok].astype(np.float64).sum(), "sum_peak_snr": snr[ok].astype(np.float64).sum(),
            "sum_psf_width": sg[ok].astype(np.float64).sum(), "sum_background": bg[ok].astype(np.float64).sum()}
    return int(ok.sum()), hists, sums


def test_histograms_and_sums_match_numpy():
    loc = _rows(20000, 1)
    p = Q.default_params(ImageWidth=256, ImageHigh=256)
    st = O.loc_stats(p, loc)
    valid, hists, sums = _numpy_hist(loc)
    assert st["valid"] == valid and valid > 1000
    for k, v in hists.items():
        np.testing.assert_array_equal(st[k], v, err_msg=k)
    for k, v in sums.items():
        assert st[k] == pytest.approx(v, rel=1e-12), k
    assert st["mean_peak_snr"] == pytest.approx(sums["sum_peak_snr"] / valid, rel=1e-6)
    assert st["mean_psf_width_ctl"] == pytest.approx(-st["mean_psf_width"] * st["mean_peak_snr"], rel=1e-6)
    # photon mean: centre of mass of the bins at >= 10% of the highest, in photons
    h = hists["hist_total_photon"].astype(np.float64)
    sel = h >= 0.1 * h.max()
    want = 25.0 * (h[sel] * (np.nonzero(sel)[0] + 0.5)).sum() / h[sel].sum()
    assert st["mean_total_photon"] == pytest.approx(want, rel=1e-5)


def _density_rows(n, seed, nframes, W, H):
    rng = np.random.default_rng(seed)
    loc = np.zeros((n, 12), np.float32)
    loc[:, 1] = rng.uniform(0, W, n)
    loc[:, 2] = rng.uniform(0, H, n)
    loc[:, 11] = np.sort(rng.integers(1, nframes + 1, n))
    return loc


def test_neighbour_counts_match_kdtree():
    W = H = 300
    loc = _density_rows(6000, 2, 8, W, H)
    p = Q.default_params(ImageWidth=W, ImageHigh=H, LocType=Q.LOCTYPE_GS2D)
    st = O.loc_stats(p, loc)
    assert (st["density_first_frame"], st["density_last_frame"]) == (2, 7)
    n0 = n1 = tot = 0
    for f in range(2, 8):
        x, y = loc[:, 1].astype(int), loc[:, 2].astype(int)
        m = (loc[:, 11] == f) & (x >= 5) & (x < W - 5) & (y >= 5) & (y < H - 5)
        pts = loc[m][:, 1:3].astype(np.float64)
        tree = cKDTree(pts)
        cnt = np.array([len(tree.query_ball_point(q, 10.0)) - 1 for q in pts])
        n0 += int((cnt == 0).sum())
        n1 += int((cnt == 1).sum())
        tot += len(pts)
    assert (st["density_total"], st["density_n0"], st["density_n1"]) == (tot, n0, n1)
    assert st["loc_density_2d"] > 0


def test_neighbour_counts_on_lattices():
    p = Q.default_params(ImageWidth=400, ImageHigh=400, LocType=Q.LOCTYPE_GS2D, PixelSize=100.0)
    # 15 px grid: nobody within 10 px
    g = np.arange(20, 380, 15, dtype=np.float32)
    xx, yy = np.meshgrid(g, g)
    loc = np.zeros((xx.size, 12), np.float32)
    loc[:, 1], loc[:, 2], loc[:, 11] = xx.ravel(), yy.ravel(), 1
    st = O.loc_stats(p, loc)
    assert (st["density_first_frame"], st["density_last_frame"]) == (1, 1)
    assert st["density_total"] == xx.size and st["density_n0"] == xx.size and st["density_n1"] == 0
    # a row of points 7 px apart: the two ends have one neighbour, the rest two
    line = np.zeros((40, 12), np.float32)
    line[:, 1], line[:, 2], line[:, 11] = 20 + 7 * np.arange(40), 100.0, 3
    st = O.loc_stats(p, line)
    assert st["density_total"] == 40 and st["density_n0"] == 0 and st["density_n1"] == 2


def test_frame_range_rule():
    p = Q.default_params(ImageWidth=200, ImageHigh=200)
    for frames, want in [((3, 4, 5), (3, 5)), ((1, 2, 3, 4), (2, 3)), ((7,), (7, 7))]:
        loc = np.zeros((len(frames) * 5, 12), np.float32)
        loc[:, 1:3] = 50
        loc[:, 11] = np.repeat(frames, 5)
        st = O.loc_stats(p, loc)
        assert (st["density_first_frame"], st["density_last_frame"]) == want


def test_activation_density_formula():
    for r0, r1, rad in [(0.6, 0.25, 1.0), (0.9, 0.05, 2.0), (0.0005, 0.3, 1.0), (1.5, 2.0, 1.0)]:
        c0, c1 = min(max(r0, 0.001), 1.0), min(max(r1, 0.001), 1.0)
        a = np.pi * 0 + 3.14159 * rad * rad
        want = c0 * (-np.log(c0) / a) + (1 - c0) * (-np.log(c1 / a) / a)
        assert O.activation_density(r0, r1, rad) == pytest.approx(want, rel=1e-5)
    # the density estimate inverts the Poisson neighbour law: for density rho the fraction with
    # no neighbour within r is exp(-rho pi r^2)
    rho, rad = 0.8, 1.0
    r0 = np.exp(-rho * 3.14159 * rad * rad)
    r1 = rho * 3.14159 * rad * rad * r0
    d = O.activation_density(r0, r1, rad)
    assert d == pytest.approx(rho, rel=0.2)
