"""GPU colour-coded 3D image (qcs_render_rgb_dev, qcs_session_get_sr_image3d) vs the oracle: on the
device's own intensity and z planes the brightness scale and every byte agree exactly, in all four
encodings and both colour modes."""
import numpy as np
import pytest
import torch

import oracle_lib as O
import qcstorm as Q

pytestmark = pytest.mark.gpu


def _rows(n, seed, W, H):
    rng = np.random.default_rng(seed)
    loc = np.zeros((n, 12), np.float32)
    loc[:, 0] = rng.uniform(100, 3000, n)
    loc[:, 1] = rng.uniform(3, W - 3, n)
    loc[:, 2] = rng.uniform(3, H - 3, n)
    loc[:, 3] = rng.uniform(-700, 700, n)
    loc[:, 4:6] = rng.uniform(1.0, 1.6, (n, 2))
    loc[:, 6] = loc[:, 0] * 6
    loc[:, 7] = rng.uniform(10, 90, n)
    loc[:, 8] = rng.uniform(3, 40, n)
    loc[:, 9] = rng.uniform(4, 35, n)
    loc[:, 10] = loc[:, 9]
    loc[:, 11] = rng.integers(1, 20, n)
    return loc


@pytest.mark.parametrize("cmode", [0, 1])
def test_rgb_matches_oracle(gpu, cmode):
    W, H = 160, 120
    params = Q.default_params(ImageWidth=W, ImageHigh=H, PixelZoom=8.0, LocType=Q.LOCTYPE_AS3D, ColorMode_3D=cmode,
                              MinZDepth=-500.0, MaxZDepth=500.0)
    loc = _rows(30000, 5 + cmode, W, H)
    ctx = Q.Context(params)
    img = torch.zeros((params.SRImageHigh, params.SRImageWidth), dtype=torch.float32, device="cuda")
    zimg = torch.zeros_like(img)
    ctx.render_dev(torch.from_numpy(loc).cuda(), loc.shape[0], img, zimg)
    torch.cuda.synchronize()
    h_img, h_z = img.cpu().numpy(), zimg.cpu().numpy()
    want_scale, _ = O.display_max_intensity(h_img)
    for mode, nb in [(Q.RGB_BGR, 3), (Q.RGB_RGB, 3), (Q.RGB_BGRA, 4), (Q.RGB_RGBA, 4)]:
        out = torch.zeros(img.numel() * nb, dtype=torch.uint8, device="cuda")
        scale = ctx.render_rgb_dev(img, zimg, mode, out)
        torch.cuda.synchronize()
        got = out.cpu().numpy().reshape(img.shape + (nb,))
        want = O.display_rgb(params, h_img, h_z, want_scale, mode)
        assert scale == want_scale
        assert np.array_equal(got, want), (mode, int((got != want).any(-1).sum()))
    assert (h_img > 0).mean() > 0.2   # a populated image


def test_session_image3d(gpu):
    W, H, nf = 128, 96, 4
    sp = Q.default_sim(width=W, height=H, density_per_px=0.004, astig=1)
    frames = Q.sim_frames_host(sp, 77, 1, nf)
    params = Q.default_params(QE=1.0, ImageWidth=W, ImageHigh=H, LocType=Q.LOCTYPE_AS3D, ROISize=9, PixelZoom=5.0,
                              p1_XGY=-111.11, p1_XLY=-111.11)
    Q.lib().qcs_params_update_sr_size(Q.C.byref(params))
    s = Q.Session(params)
    s.feed(frames)
    s.finish()
    s.wait()
    im3 = s.sr_image3d()
    sr = s.sr_image()
    s.close()
    assert im3.shape == (params.SRImageHigh, params.SRImageWidth, 4)
    assert (im3[..., 3] == 255).all()
    assert not im3[sr == 0][..., :3].any()          # black background
    # value = trunc(intensity * 120 / scale): pixels under scale / 120 are black, brighter ones not
    scale, _ = O.display_max_intensity(sr)
    v = sr * 120.0 / scale
    lit = im3[..., :3].any(-1)
    assert lit[v >= 1.01].all() and not lit[(v < 0.99)].any() and lit.sum() > 100
    with pytest.raises(Q.QcsError):                  # 2D sessions have no z plane
        s2 = Q.Session(Q.default_params(ImageWidth=64, ImageHigh=64))
        try:
            s2.sr_image3d()
        finally:
            s2.close()
