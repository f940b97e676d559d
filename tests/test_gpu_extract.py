"""GPU ROI extraction (filters, detection, classification, cut, WLE) vs the oracle: bit-exact."""
import numpy as np
import pytest
import torch

import oracle_lib as O
import qcstorm as Q

pytestmark = pytest.mark.gpu

CASES = [
    # W, H, nframes, B, loc_type, multi, R, density, cap
    (256, 256, 3, 1, 0, 0, 7, 0.003, 100000),
    (256, 192, 4, 2, 0, 1, 7, 0.02, 100000),     # 2D multi: unsmoothed detection + classification, batched seams
    (200, 150, 3, 1, 1, 0, 9, 0.003, 100000),    # AS3D, width not a multiple of 32/64
    (320, 256, 2, 1, 0, 1, 11, 0.03, 100000),
    (256, 256, 2, 1, 0, 0, 7, 0.01, 50),         # capacity overflow: first 50 in raster order
    (2048, 2048, 16, 2, 0, 0, 7, 0.003, 400000), # full-size frames: 128-row walks, 37 strips, seams
    (1000, 700, 3, 3, 0, 1, 9, 0.01, 100000),    # W not a multiple of 56 or 32, 3-frame batch
]


@pytest.mark.parametrize("W,H,nf,B,lt,mu,R,dens,cap", CASES)
def test_extract_bit_exact(gpu, W, H, nf, B, lt, mu, R, dens, cap):
    sp = Q.default_sim(width=W, height=H, density_per_px=dens, astig=int(lt == 1))
    frames = Q.sim_frames_host(sp, 31 + W + nf, 5, nf)
    params = Q.default_params(QE=1.0, ImageWidth=W, ImageHigh=H, LocType=lt, MultiEmitterFitEn=mu, ROISize=R,
                              BatchedImageNum=B)
    ref_rois, ref_wle, ref_found, _ = O.extract(params, frames, start_frame=5, cap=cap)
    ctx = Q.Context(params)
    got_rois, got_wle, got_found = ctx.extract(frames, start_frame=5, cap=cap)
    print("case", W, H, nf, B, lt, mu, R, "found", ref_found, got_found, "kept", len(ref_rois), len(got_rois),
          "multi-type", int((ref_rois[:, R * R + 4] == 1).sum()) if len(ref_rois) else 0)
    assert got_found == ref_found
    assert got_rois.shape == ref_rois.shape
    assert np.array_equal(got_rois, ref_rois)
    bad = np.argwhere(got_wle.view(np.uint32) != ref_wle.view(np.uint32))
    if len(bad):
        i = bad[0][0]
        print("wle row", i, got_wle[i], ref_wle[i])
    assert len(bad) == 0


def test_extract_device_api_matches_host_api(gpu):
    W = H = 256
    sp = Q.default_sim(width=W, height=H, density_per_px=0.004)
    frames = Q.sim_frames_host(sp, 77, 1, 3)
    params = Q.default_params(QE=1.0, ImageWidth=W, ImageHigh=H)
    ctx = Q.Context(params)
    h_rois, h_wle, h_found = ctx.extract(frames, start_frame=1, cap=20000)
    d_frames = torch.from_numpy(frames.view(np.int16)).cuda()
    d_rois = torch.zeros((20000, Q.roi_words(7)), dtype=torch.int16, device="cuda")
    d_wle = torch.zeros((20000, 6), dtype=torch.float32, device="cuda")
    n, found = ctx.extract_dev(d_frames, 3, 1, d_rois, d_wle, 20000)
    torch.cuda.synchronize()
    assert n == len(h_rois) and found == h_found
    assert np.array_equal(d_rois[:n].cpu().numpy().view(np.uint16), h_rois)
    assert np.array_equal(d_wle[:n].cpu().numpy(), h_wle)
