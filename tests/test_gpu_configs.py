"""Parity at the shapes of the astigmatic-3D and high-density multi-emitter workloads: whole-frame
extraction, fitting and frame-sorted compaction (qcs_loc_batch_sync) on 1024 x 1024 frames
against the oracle pipeline, bit for bit, plus the accuracy of the astigmatic z lookup against the
simulated truth."""
import numpy as np
import pytest

import oracle_lib as O
import qcstorm as Q

pytestmark = pytest.mark.gpu


def _oracle_rows(params, frames, first_frame):
    p = Q.default_params(**{k: getattr(params, k) for k, _ in Q.QcsParams._fields_})
    p.BatchedImageNum = frames.shape[0]
    rois, wles, _, _ = O.extract(p, frames, start_frame=first_frame, cap=400000)
    loc, _ = O.fit(params, rois, wles)
    return O.compact_sorted(loc, first_frame, first_frame + frames.shape[0] - 1)


def test_astigmatic_3d_1024_frames(gpu):
    """1024 x 1024 frames, 0.3 emitters/um^2, z uniform in +-500 nm, sigma_x/y = 1.2 sqrt(1 + ((z -+ 250) / 400)^2),
    photons 1000-8000, background 20-80, 9 x 9 ROIs, z = -111.11 (sigma_x^2 - sigma_y^2)."""
    W = H = 1024
    sp = Q.default_sim(width=W, height=H, density_per_px=0.003, astig=1, photons_min=1000, photons_max=8000,
                       bg_min=20, bg_max=80)
    first, nf = 11, 3
    frames = Q.sim_frames_host(sp, 1002, first, nf)
    params = Q.default_params(QE=1.0, ImageWidth=W, ImageHigh=H, LocType=Q.LOCTYPE_AS3D, ROISize=9,
                              p1_XGY=-111.11, p1_XLY=-111.11, ZDepthCorrFactor=1.0)
    ctx = Q.Context(params)
    got = ctx.loc_batch_sync(frames, start_frame=first)
    ref = _oracle_rows(params, frames, first)
    print("astig rows", got.shape[0], ref.shape[0])
    assert got.shape == ref.shape and ref.shape[0] > 5000
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
    # z against the simulated emitters (nearest within 0.5 px, same frame)
    dz, dxy = [], []
    for f in range(first, first + nf):
        t = Q.frame_emitters(sp, 1002, f)
        rows = got[got[:, 11] == f]
        d2 = (rows[:, None, 1] - t[None, :, 1]) ** 2 + (rows[:, None, 2] - t[None, :, 2]) ** 2
        j = d2.argmin(1)
        ok = d2[np.arange(len(rows)), j] < 0.25
        dz.append(rows[ok, 3] - t[j[ok], 3])
        dxy.append(np.sqrt(d2[np.arange(len(rows)), j][ok]))
    dz, dxy = np.concatenate(dz), np.concatenate(dxy)
    inner = np.abs(np.concatenate([Q.frame_emitters(sp, 1002, f)[:, 3] for f in range(first, first + nf)])).size
    rmse_z = float(np.sqrt(np.mean(dz[np.abs(dz) < 300] ** 2)))
    print("matched", dz.size, "of", inner, "z rmse nm", rmse_z, "median |dz|", float(np.median(np.abs(dz))),
          "xy rmse px", float(np.sqrt(np.mean(dxy ** 2))))
    assert dz.size > 0.6 * inner
    assert np.median(np.abs(dz)) < 60.0 and rmse_z < 120.0


def test_high_density_multi_emitter_1024_frame(gpu):
    """1024 x 1024, 3 emitters/um^2 (about 31k per frame), multi-emitter cascade, 7 x 7 ROIs."""
    W = H = 1024
    sp = Q.default_sim(width=W, height=H, density_per_px=0.03)
    first = 21
    frames = Q.sim_frames_host(sp, 1003, first, 1)
    params = Q.default_params(QE=1.0, ImageWidth=W, ImageHigh=H, MultiEmitterFitEn=1, ROISize=7)
    ctx = Q.Context(params)
    got = ctx.loc_batch_sync(frames, start_frame=first)
    st = ctx.stats()
    ref = _oracle_rows(params, frames, first)
    print("dense rows", got.shape[0], ref.shape[0], "2E", st["n_2e"], "3E", st["n_3e"])
    assert got.shape == ref.shape and ref.shape[0] > 10000
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
    assert st["n_2e"] > 0
