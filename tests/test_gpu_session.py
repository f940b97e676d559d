"""Online session (lm_* flow) and the synchronous batch call vs the oracle pipeline.

The oracle pipeline restates the reference's online loop (IJ/OnlineLocalizationLD.cpp:128-300):
each fed group of frames is one batched image for extraction, ROIs accumulate until PointNumTh
(or the end), then fit + CRLB, frame-sorted removal of rejected rows over [first, end] and
rendering.  Rows are compared bit-exactly; the SR image (float atomics) within 1e-5 relative.
"""
import numpy as np
import pytest

import oracle_lib as O
import qcstorm as Q

pytestmark = pytest.mark.gpu


def _oracle_pipeline(params, groups, first_frame=1):
    rois, wles = [], []
    f = first_frame
    for g in groups:
        p = Q.default_params(**{k: getattr(params, k) for k, _ in Q.QcsParams._fields_})
        p.BatchedImageNum = g.shape[0]
        r, w, _, _ = O.extract(p, g, start_frame=f, cap=200000)
        rois.append(r)
        wles.append(w)
        f += g.shape[0]
    rois, wles = np.concatenate(rois), np.concatenate(wles)
    loc, _ = O.fit(params, rois, wles)
    return O.compact_sorted(loc, first_frame, f - 1)


@pytest.mark.parametrize("lt,mu,R,group", [(0, 0, 7, 1), (0, 1, 7, 2), (1, 0, 9, 3)])
def test_session_matches_oracle_pipeline(gpu, lt, mu, R, group):
    W, H, nf = 256, 192, 6
    sp = Q.default_sim(width=W, height=H, density_per_px=0.006 if mu else 0.003, astig=int(lt == 1))
    frames = Q.sim_frames_host(sp, 500 + R + group, 1, nf)
    params = Q.default_params(QE=1.0, ImageWidth=W, ImageHigh=H, LocType=lt, MultiEmitterFitEn=mu, ROISize=R,
                              PixelZoom=5.0, p1_XGY=-111.11, p1_XLY=-111.11, ZDepthCorrFactor=1.0)
    Q.lib().qcs_params_update_sr_size(Q.C.byref(params))
    groups = [frames[i:i + group] for i in range(0, nf, group)]
    s = Q.Session(params)
    for g in groups:
        s.feed(g)
    s.finish()
    s.wait()
    got = s.locs()
    img = s.sr_image()
    mx = s.max_disp_val()
    s.close()
    ref = _oracle_pipeline(params, groups)
    print("case", lt, mu, R, group, "rows", got.shape[0], ref.shape[0])
    assert got.shape == ref.shape and ref.shape[0] > 50
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
    rimg, _ = O.render(params, ref, with_z=(lt == 1))
    np.testing.assert_allclose(img, rimg, rtol=1e-5, atol=1e-5 * max(float(rimg.max()), 1.0))
    assert mx == max(int(img.max()), 1)


def _oracle_online(params, groups, threshold):
    """The online loop with flushes (IJ/OnlineLocalizationLD.cpp:195-300): ROIs of fed groups
    accumulate until `threshold` (or the end); each flush fits them, removes rejected rows with
    frame sorting, merges consecutive frames (held-back group; the end flush gives everything
    back) and removes the merged-away rows.  Returns the rows of every flush."""
    cs = O.Consec(params) if params.ConsecFitEn else None
    out, rois, wles = [], [], []
    f, first = 1, 1

    def flush(end):
        nonlocal rois, wles, first
        if rois:
            loc, _ = O.fit(params, np.concatenate(rois), np.concatenate(wles))
            rows = O.compact_sorted(loc, first, f - 1)
        else:
            rows = np.zeros((0, 12), np.float32)
        if cs is not None:
            if not rois and not (end and cs.held()):
                return
            merged, ratio = cs.fit(rows, end)
            rows = O.compact(merged)
        elif not rois:
            return
        out.append(rows)
        rois, wles = [], []

    for g in groups:
        p = Q.default_params(**{k: getattr(params, k) for k, _ in Q.QcsParams._fields_})
        p.BatchedImageNum = g.shape[0]
        r, w, _, _ = O.extract(p, g, start_frame=f, cap=200000)
        if not rois:
            first = f
        rois.append(r)
        wles.append(w)
        f += g.shape[0]
        if sum(x.shape[0] for x in rois) >= threshold:
            flush(False)
    flush(True)
    return out


@pytest.mark.parametrize("consec,th_nm", [(1, 100.0), (1, 160.0), (0, 100.0)])
def test_session_flushes_consec_and_result_file(gpu, tmp_path, consec, th_nm):
    W, H, nf = 128, 128, 30
    sp = Q.default_sim(width=W, height=H, density_per_px=0.004)
    frames = Q.sim_frames_host(sp, 900 + consec, 1, nf)
    params = Q.default_params(QE=1.0, ImageWidth=W, ImageHigh=H, ConsecFitEn=consec, ConsecFit_DistanceTh_nm=th_nm,
                              FlushROINum=300)
    groups = [frames[i:i + 2] for i in range(0, nf, 2)]
    s = Q.Session(params)
    path = tmp_path / Q.result_file_name(params, "movie.tif")
    s.set_output(str(path))
    for g in groups:
        s.feed(g)
    s.finish()
    s.wait()
    got = s.locs()
    st, ratio = s.stats()
    s.close()
    flushes = _oracle_online(params, groups, 300)
    ref = np.concatenate(flushes)
    print("consec", consec, th_nm, "flushes", len(flushes), "rows", got.shape[0], ref.shape[0])
    assert len(flushes) >= 3
    assert got.shape == ref.shape and ref.shape[0] > 100
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
    on_disk = np.fromfile(path, dtype="<f4").reshape(-1, 12)
    assert np.array_equal(on_disk.view(np.uint32), got.view(np.uint32))
    rst = O.loc_stats(params, flushes[-1])
    for k in ("valid", "density_total", "density_n0", "density_n1"):
        assert st[k] == rst[k], k
    if consec:
        assert ratio.sum() == pytest.approx(1.0, abs=1e-6) or ratio.sum() == 0.0


def test_session_output_errors(gpu, tmp_path):
    params = Q.default_params(QE=1.0, ImageWidth=64, ImageHigh=64)
    s = Q.Session(params)
    with pytest.raises(Q.QcsError):
        s.set_output(str(tmp_path / "no_such_dir" / "x.txt"))
    s.feed(np.zeros((1, 64, 64), np.uint16) + 100)
    with pytest.raises(Q.QcsError):
        s.set_output(str(tmp_path / "late.txt"))
    s.close()


def test_loc_batch_sync_matches_oracle(gpu):
    W, H = 256, 256
    sp = Q.default_sim(width=W, height=H, density_per_px=0.003)
    frames = Q.sim_frames_host(sp, 71, 5, 3)
    params = Q.default_params(QE=1.0, ImageWidth=W, ImageHigh=H)
    ctx = Q.Context(params)
    got = ctx.loc_batch_sync(frames, start_frame=5)
    ref = _oracle_pipeline(params, [frames], first_frame=5)
    assert got.shape == ref.shape and ref.shape[0] > 50
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))


def test_loc_batch_sync_info_matches_oracle(gpu):
    """lm_LocBatchedImg's return array: statistics of the batch's rows, per call."""
    W, H = 256, 256
    sp = Q.default_sim(width=W, height=H, density_per_px=0.006)
    frames = Q.sim_frames_host(sp, 72, 1, 6)
    params = Q.default_params(QE=1.0, ImageWidth=W, ImageHigh=H)
    ctx = Q.Context(params)
    for f0, nf in ((1, 6), (3, 2)):
        rows, info, st = ctx.loc_batch_sync_info(frames[f0 - 1:f0 - 1 + nf], start_frame=f0)
        ref = _oracle_pipeline(params, [frames[f0 - 1:f0 - 1 + nf]], first_frame=f0)
        assert np.array_equal(rows.view(np.uint32), ref.view(np.uint32))
        rst = O.loc_stats(params, ref)
        for k in ("valid", "density_total", "density_n0", "density_n1", "density_first_frame", "density_last_frame"):
            assert st[k] == rst[k], k
        np.testing.assert_array_equal(st["hist_peak_snr"], rst["hist_peak_snr"])
        snr, psf = rst["mean_peak_snr"], rst["mean_psf_width"]
        want = [snr, -psf, snr - psf, snr * -psf, ref.shape[0], rst["loc_density_2d"], 0, 0]
        np.testing.assert_allclose(info, np.array(want, np.float32), rtol=1e-6, atol=1e-30)
    assert info[5] == 0 or st["density_total"] > 50


def test_session_rejects_unsupported_modes(gpu):
    with pytest.raises(Q.QcsError):
        Q.Session(Q.default_params(ImageWidth=64, ImageHigh=64, FlushROINum=-1))
    with pytest.raises(Q.QcsError):
        Q.Session(Q.default_params(ImageWidth=64, ImageHigh=64, LocType=Q.LOCTYPE_DH3D))
