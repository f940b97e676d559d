"""Pin the oracle's compaction, rendering, detector classification, ROI cut and AS3D z lookup
against numpy restatements and closed forms."""
import numpy as np

import oracle_lib as O
import qcstorm as Q


def _rows(n, seed, nframes=20):
    rng = np.random.default_rng(seed)
    loc = np.zeros((n, 12), np.float32)
    loc[:, 0] = rng.choice([np.nan, 0.0, 0.5, 1.0, 1.5, 1.99, 2.0, 3.0, 250.0, 1200.0], n)
    loc[:, 1:11] = rng.uniform(0, 50, (n, 10))
    loc[:, 11] = rng.integers(0, nframes + 2, n)
    return loc


def _trunc_sat(v):
    out = np.zeros(v.shape, np.int64)
    ok = ~np.isnan(v)
    out[ok] = np.trunc(v[ok]).astype(np.int64)
    return out


def test_compact_sorted_is_stable_frame_sort_of_int_ppho_above_one():
    # ZeroLocalizationsRemove.cu:171 keeps (int)PPho > 1 (so PPho in (1, 2) is dropped) per frame
    loc = _rows(4000, 11)
    first, end = 3, 15
    pp, fr = _trunc_sat(loc[:, 0]), _trunc_sat(loc[:, 11])
    keep = (pp > 1) & (fr >= first) & (fr <= end)
    idx = np.nonzero(keep)[0]
    idx = idx[np.argsort(fr[idx], kind="stable")]
    got = O.compact_sorted(loc, first, end)
    assert np.array_equal(got.view(np.uint32), loc[idx].view(np.uint32))
    assert not np.any((loc[idx, 0] > 1.0) & (loc[idx, 0] < 2.0))


def test_compact_unsorted_keeps_ppho_above_one_in_order():
    loc = _rows(4000, 12)
    keep = loc[:, 0] > 1.0   # NaN compares false
    got = O.compact(loc)
    assert np.array_equal(got.view(np.uint32), loc[keep].view(np.uint32))


def _render_params(**kw):
    p = Q.default_params(ImageWidth=64, ImageHigh=48, PixelZoom=10.0, PixelSize=100.0, SNR_th=5.0, **kw)
    Q.lib().qcs_params_update_sr_size(Q.C.byref(p))
    return p


def test_render_single_localization_closed_form():
    # one row: the 7x7 window at floor(X*zoom)-3.. holds PPho * exp(-((c - fx)^2 + (r - fy)^2) / (2 s^2)),
    # fx = frac(X*zoom) + 3, s = max(CRLBx, px/zoom/2) / (px/zoom)  (ImageRender_2d.cu:25-155)
    p = _render_params()
    row = np.zeros((1, 12), np.float32)
    row[0, 0], row[0, 1], row[0, 2], row[0, 8], row[0, 9], row[0, 11] = 1000.0, 20.37, 17.81, 12.0, 23.0, 1
    img, _ = O.render(p, row)
    rx, ry = np.float32(row[0, 1]) * np.float32(10.0), np.float32(row[0, 2]) * np.float32(10.0)
    ix, iy = int(rx), int(ry)
    s = 23.0 / 10.0
    fx, fy = float(rx) - ix + 3.0, float(ry) - iy + 3.0
    yy, xx = np.mgrid[0:7, 0:7]
    ref = 1000.0 * np.exp(-((xx - fx) ** 2 + (yy - fy) ** 2) / (2 * s * s))
    np.testing.assert_allclose(img[iy - 3:iy + 4, ix - 3:ix + 4], ref, rtol=2e-6, atol=1e-3)
    assert np.count_nonzero(img) == 49
    np.testing.assert_allclose(img.sum(), ref.sum(), rtol=1e-6)


def test_render_skips_invalid_rows():
    p = _render_params()
    base = np.array([1000.0, 20.0, 20.0, 0, 1.2, 1.2, 0, 30.0, 12.0, 15.0, 15.0, 1], np.float32)
    bad = np.tile(base, (6, 1))
    bad[0, 8] = 4.0            # SNR below SNR_th
    bad[1, 0] = 1.0            # PPho <= 1
    bad[2, 9] = np.nan         # NaN precision
    bad[3, 9] = -1.0           # non-positive precision
    bad[4, 1] = 0.5            # inside the 7-pixel SR border
    bad[5, 9] = 31.0           # 2D precision above 30 nm: weight 0
    img, _ = O.render(p, bad)
    assert float(np.abs(img).sum()) == 0.0
    ok, _ = O.render(p, base[None, :])
    assert ok.sum() > 0


def test_detector_classification_counts_marks_outside_the_centre():
    # gpuDetectedROIClasify (LDROIExtraction_ROIExtraction.cu:181-243): multi when the R x R window
    # holds a mark outside the central 3 x 3
    rng = np.random.default_rng(3)
    w, h, R = 60, 50, 7
    mark = (rng.random((h, w)) < 0.02).astype(np.uint8)
    ys, xs = rng.integers(3, h - 3, 300), rng.integers(3, w - 3, 300)
    xy = np.stack([xs, ys], 1).astype(np.int32).ravel()
    ft = np.zeros(300, np.int32)
    O.lib().orc_classify_rois(O.p(mark), O.p(xy), 300, 1, R, w, O.p(ft))
    for i in range(300):
        x, y = xs[i], ys[i]
        d = int(mark[y - 3:y + 4, x - 3:x + 4].sum()) - int(mark[y - 1:y + 2, x - 1:x + 2].sum())
        assert ft[i] == (1 if d > 0 else 0)
    O.lib().orc_classify_rois(O.p(mark), O.p(xy), 300, 0, R, w, O.p(ft))
    assert not ft.any()


def test_roi_cut_is_the_raw_window_and_header():
    rng = np.random.default_rng(4)
    w, h, R, B = 40, 30, 9, 2
    raw = rng.integers(0, 65536, (B * h, w)).astype(np.uint16)
    xs, ys = np.array([4, 20, 35]), np.array([4, 33, 55])
    xy = np.stack([xs, ys], 1).astype(np.int32).ravel()
    ft = np.array([0, 1, 0], np.int32)
    out = np.zeros((3, R * (R + 1)), np.uint16)
    O.lib().orc_cut_rois(O.p(raw), O.p(xy), O.p(ft), 3, R, w, B * h, h, 70000, O.p(out))
    for i in range(3):
        x, y = xs[i], ys[i]
        assert np.array_equal(out[i, :R * R].reshape(R, R), raw[y - 4:y + 5, x - 4:x + 5])
        fr = 70000 + y // h
        assert list(out[i, R * R:R * R + 5]) == [x, y % h, fr % 65536, fr // 65536, ft[i]]
        assert not out[i, R * R + 5:].any()


def test_as3d_z_lookup_formula_and_physics():
    # bfgsMLE_AS3D_core.h:266-279: z = sum p_k d^k * ZDepthCorrFactor with d = sigx^2 - sigy^2.
    # For sigx/y(z) = s0 sqrt(1 + ((z -+ c)/d)^2): sigx^2 - sigy^2 = -4 s0^2 c z / d^2, so
    # p1 = -d^2 / (4 s0^2 c) = -111.11 nm/px^2 (s0 = 1.2 px, c = 250 nm, d = 400 nm) inverts it.
    sp = Q.default_sim(width=9, height=9, astig=1, photons_min=3000.0, photons_max=8000.0, bg_min=20.0,
                       bg_max=40.0, center_jitter=0.5)
    rois, truth = Q.sim_rois_host(sp, 77, 800)
    p = Q.default_params(QE=1.0, ROISize=9, LocType=Q.LOCTYPE_AS3D, p1_XGY=-111.11, p1_XLY=-111.11,
                         ZDepthCorrFactor=1.0)
    loc, _ = O.fit(p, rois, O.wle(rois, 9, 1))
    loc = loc[:800]
    ok = loc[:, 0] > 0
    assert ok.sum() > 600
    f32 = np.float32
    sx, sy = loc[ok, 4].astype(f32), loc[ok, 5].astype(f32)
    d = (sx * sx).astype(f32) - (sy * sy).astype(f32)
    assert np.array_equal(loc[ok, 3], (f32(-111.11) * d).astype(f32))
    z_true = truth[:800, 0, 3][ok]
    err = loc[ok, 3] - z_true
    assert np.corrcoef(loc[ok, 3], z_true)[0, 1] > 0.9
    assert abs(np.median(err)) < 40.0
