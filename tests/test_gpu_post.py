"""GPU frame-sorted compaction (exact) and splat rendering (float-atomic tolerance) vs the oracle."""
import numpy as np
import pytest
import torch

import oracle_lib as O
import qcstorm as Q

pytestmark = pytest.mark.gpu


def _rows(n, seed, nframes=20, zrange=0.0):
    rng = np.random.default_rng(seed)
    loc = np.zeros((n, 12), np.float32)
    loc[:, 0] = rng.choice([0.0, 0.5, 1.5, 2.0, 3.0, 250.0, 1200.0], n)      # includes PPho in (1, 2)
    loc[:, 1] = rng.uniform(3, 197, n)
    loc[:, 2] = rng.uniform(3, 147, n)
    loc[:, 3] = rng.uniform(-zrange, zrange, n)
    loc[:, 4:6] = rng.uniform(1.0, 1.6, (n, 2))
    loc[:, 6] = loc[:, 0] * 10
    loc[:, 7] = rng.uniform(10, 90, n)
    loc[:, 8] = rng.uniform(0, 40, n)
    loc[:, 9] = rng.choice([np.nan, -1.0, 5.0, 12.0, 25.0, 40.0], n)
    loc[:, 10] = rng.uniform(4, 35, n)
    loc[:, 11] = rng.integers(1, nframes + 1, n)
    return loc


def test_compact_sorted_exact(gpu):
    loc = _rows(5000, 1)
    ctx = Q.Context(Q.default_params())
    d = torch.from_numpy(loc).cuda()
    out = torch.zeros_like(d)
    m = ctx.compact_sorted_dev(d, loc.shape[0], 3, 17, out)
    torch.cuda.synchronize()
    ref = O.compact_sorted(loc, 3, 17)
    assert m == len(ref) and np.array_equal(out[:m].cpu().numpy().view(np.uint32), ref.view(np.uint32))
    assert np.all(np.diff(ref[:, 11]) >= 0) and np.all(ref[:, 0] >= 2.0)


def test_compact_unsorted_exact(gpu):
    loc = _rows(5000, 2)
    ctx = Q.Context(Q.default_params())
    d = torch.from_numpy(loc).cuda()
    out = torch.zeros_like(d)
    m = ctx.compact_dev(d, loc.shape[0], out)
    torch.cuda.synchronize()
    ref = O.compact(loc)
    assert m == len(ref) and np.array_equal(out[:m].cpu().numpy().view(np.uint32), ref.view(np.uint32))


@pytest.mark.parametrize("lt,mode", [(0, 0), (0, 1), (0, 2), (0, 3), (1, 0)])
def test_render_matches_oracle(gpu, lt, mode):
    loc = _rows(20000, 3 + mode, zrange=600.0 if lt else 0.0)
    params = Q.default_params(ImageWidth=200, ImageHigh=150, PixelZoom=10.0, LocType=lt, RenderingMode=mode,
                              FixedLocPrec=15.0)
    ref, zref = O.render(params, loc, with_z=(lt == 1))
    ctx = Q.Context(params)
    d = torch.from_numpy(loc).cuda()
    img = torch.zeros((params.SRImageHigh, params.SRImageWidth), dtype=torch.float32, device="cuda")
    zimg = torch.zeros_like(img) if lt == 1 else None
    ctx.render_dev(d, loc.shape[0], img, zimg)
    torch.cuda.synchronize()
    got = img.cpu().numpy()
    # same PSF values, float atomics in arbitrary order: relative agreement to a few ulp of the sum
    scale = np.maximum(np.abs(ref), 1e-3)
    print("mode", mode, "sum", ref.sum(), "max rel", np.max(np.abs(got - ref) / scale))
    assert np.all(np.abs(got - ref) <= 1e-5 * scale + 1e-6)
    assert ref.sum() > 0
    if lt == 1:
        zg = zimg.cpu().numpy()
        assert np.allclose(zg, zref, rtol=1e-4, atol=1e-2)


def test_render_accumulates_into_existing_image(gpu):
    """Rendering adds to the image (tile read-modify-write): two calls give twice one call, and
    pixels no localization touches keep their value."""
    loc = _rows(20000, 11)
    params = Q.default_params(ImageWidth=200, ImageHigh=150, PixelZoom=10.0)
    ref, _ = O.render(params, loc, with_z=False)
    ctx = Q.Context(params)
    d = torch.from_numpy(loc).cuda()
    img = torch.full((params.SRImageHigh, params.SRImageWidth), 0.25, dtype=torch.float32, device="cuda")
    ctx.render_dev(d, loc.shape[0], img, None)
    ctx.render_dev(d, loc.shape[0], img, None)
    torch.cuda.synchronize()
    got = img.cpu().numpy()
    want = 0.25 + 2.0 * ref
    scale = np.maximum(np.abs(want), 1e-3)
    assert np.all(np.abs(got - want) <= 2e-5 * scale + 1e-6)
    assert np.all(got[ref == 0] == np.float32(0.25))


@pytest.mark.parametrize("lt,unaligned", [(0, False), (0, True), (1, False)])
def test_render_fresh_replaces_the_planes(gpu, lt, unaligned):
    """qcs_render_fresh_dev equals clearing the planes and rendering: pixels no localization touches
    end at zero whatever the planes held, the others match the oracle's image.  An image that
    starts 4 bytes past a 16-byte boundary takes the per-pixel path in every tile; the aligned
    one the 16-byte path inside and the per-pixel path in the partial tiles of the right and
    bottom edges (2000 x 1500 SR pixels, 64-pixel tiles)."""
    loc = _rows(20000, 21 + int(unaligned), zrange=600.0 if lt else 0.0)
    params = Q.default_params(ImageWidth=200, ImageHigh=150, PixelZoom=10.0, LocType=lt, FixedLocPrec=15.0)
    ref, zref = O.render(params, loc, with_z=(lt == 1))
    ctx = Q.Context(params)
    d = torch.from_numpy(loc).cuda()
    H, W = params.SRImageHigh, params.SRImageWidth
    off = 1 if unaligned else 0
    store = torch.full((H * W + 4,), 7.0, dtype=torch.float32, device="cuda")
    img = store[off:off + H * W].view(H, W)
    zimg = torch.full((H, W), -3.0, dtype=torch.float32, device="cuda") if lt == 1 else None
    ctx.render_fresh_dev(d, loc.shape[0], img, zimg)
    torch.cuda.synchronize()
    got = img.cpu().numpy()
    scale = np.maximum(np.abs(ref), 1e-3)
    print("unaligned", unaligned, "sum", ref.sum(), "max rel", np.max(np.abs(got - ref) / scale))
    assert np.all(np.abs(got - ref) <= 1e-5 * scale + 1e-6)
    assert np.all(got[ref == 0] == 0.0) and (ref == 0).sum() > 0
    assert float(store[:off].sum().item()) == 7.0 * off and float(store[off + H * W:].min().item()) == 7.0
    if lt == 1:
        zg = zimg.cpu().numpy()
        assert np.allclose(zg, zref, rtol=1e-4, atol=1e-2)
        assert np.all(zg[zref == 0] == 0.0)
    # the accumulating call on top of a fresh image adds the same sums again
    ctx.render_dev(d, loc.shape[0], img, zimg)
    torch.cuda.synchronize()
    got2 = img.cpu().numpy()
    assert np.all(np.abs(got2 - 2.0 * ref) <= 2e-5 * scale + 1e-6)
    # no rows: the planes are cleared
    img.fill_(7.0)
    if zimg is not None:
        zimg.fill_(-3.0)
    ctx.render_fresh_dev(d, 0, img, zimg)
    torch.cuda.synchronize()
    assert float(img.abs().max().item()) == 0.0
    if zimg is not None:
        assert float(zimg.abs().max().item()) == 0.0
