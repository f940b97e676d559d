"""Pin the oracle's extraction against closed forms and the numpy formulation."""
import ctypes as C

import numpy as np
import pytest

import extract_np as E
import oracle_lib as O
import qcstorm as Q


def taps(R=7, loc_type=0, multi=0):
    t = np.zeros(5, np.float32)
    O.lib().orc_filter_taps(R, loc_type, multi, O.p(t))
    return t


def test_filter_taps_normalised_and_symmetric():
    for R, lt, mu in [(7, 0, 0), (7, 0, 1), (9, 1, 0), (11, 0, 1), (5, 0, 0)]:
        t = taps(R, lt, mu).astype(np.float64)
        assert abs(np.outer(t, t).sum() - 1.0) < 1e-6
        assert t[0] == t[4] and t[1] == t[3] and t[2] > t[1] > t[0]
    # R = 7, 2D single: sigma = 0.5 * 1.2 -> taps proportional to exp(-k^2 / 0.72)
    t = taps(7, 0, 0).astype(np.float64)
    assert abs(t[1] / t[2] - np.exp(-1 / 0.72)) < 1e-6


def _frames(w, h, n, seed, density=0.002):
    sp = Q.default_sim(width=w, height=h, density_per_px=density)
    return Q.sim_frames_host(sp, seed, 1, n)


def test_background_removal_constant_frame_is_zero():
    img = np.full((40, 70), 1234, np.uint16)
    out = np.zeros_like(img)
    O.lib().orc_remove_background(O.p(img), O.p(out), 70, 40)
    assert not out.any()


@pytest.mark.parametrize("w,h", [(128, 96), (100, 77)])
def test_background_removal_matches_numpy(w, h):
    img = _frames(w, h, 1, 11)[0]
    out = np.zeros_like(img)
    O.lib().orc_remove_background(O.p(img), O.p(out), w, h)
    assert np.array_equal(out, E.remove_background(img))


def test_smoothing_matches_numpy_and_delta_response():
    img = _frames(96, 64, 1, 12)[0]
    t = taps()
    a = np.zeros_like(img)
    b = np.zeros_like(img)
    O.lib().orc_conv_h(O.p(img), O.p(a), 96, 64, O.p(t))
    O.lib().orc_conv_v(O.p(a), O.p(b), 96, 64, O.p(t))
    assert np.array_equal(a, E.conv_h(img, t)) and np.array_equal(b, E.conv_v(a, t))
    d = np.zeros((20, 20), np.uint16)
    d[10, 10] = 10000
    h = np.zeros_like(d)
    O.lib().orc_conv_h(O.p(d), O.p(h), 20, 20, O.p(t))
    assert list(h[10, 8:13]) == [int(np.float32(10000) * t[k]) for k in range(5)]


def test_variance_closed_form():
    rng = np.random.default_rng(4)
    img = rng.integers(0, 200, (300, 400)).astype(np.uint16)
    v = O.lib().orc_image_variance(O.p(img), 400, 300)
    S = min(min(400 // 2 - 5, 256), min(300 // 2 - 5, 256))
    cx, cy = 200, 150
    blk = lambda y0, x0: img[y0:y0 + S, x0:x0 + S].astype(np.int64)
    m = np.minimum(np.minimum(blk(cy, cx), blk(cy, cx - S)), np.minimum(blk(cy - S, cx), blk(cy - S, cx - S)))
    ref = np.float32(np.mean(m.astype(np.float64) ** 2) - np.mean(m.astype(np.float64)) ** 2)
    assert v == max(ref, np.float32(100.0))


def test_find_matches_numpy():
    img = _frames(128, 128, 1, 13, density=0.004)[0]
    bg = E.remove_background(img)
    t = taps()
    det = E.conv_v(E.conv_h(bg, t), t)
    var = O.lib().orc_image_variance(O.p(bg), 128, 128)
    mark = np.zeros(det.shape, np.uint8)
    xy = np.zeros((4096, 2), np.int32)
    n = O.lib().orc_find_rois(O.p(det), 128, 128, 128, var, 0, 0, 7, O.p(mark), O.p(xy), 4096)
    ok, ref = E.find(det, 128, var, 7, 8.5, 3.5)
    assert n == len(ref) and np.array_equal(xy[:n], ref) and np.array_equal(mark.astype(bool), ok)
    assert n > 5


def test_extract_finds_simulated_emitters():
    sp = Q.default_sim(width=256, height=256, density_per_px=0.0005, photons_min=3000, photons_max=5000)
    frames = Q.sim_frames_host(sp, 21, 1, 1)
    params = Q.default_params(QE=1.0, ImageWidth=256, ImageHigh=256)
    rois, w, found, var = O.extract(params, frames)
    tr = Q.frame_emitters(sp, 21, 1)
    inside = (tr[:, 1] > 6) & (tr[:, 1] < 250) & (tr[:, 2] > 6) & (tr[:, 2] < 250)
    cx = rois[:, 49].astype(float) + 0.5
    cy = rois[:, 50].astype(float) + 0.5
    d = np.hypot(cx[:, None] - tr[None, :, 1], cy[:, None] - tr[None, :, 2]).min(1)
    # every detection sits on a simulated emitter, and most bright emitters are detected
    assert found == len(rois) and np.all(d < 1.5) and len(rois) >= 0.8 * inside.sum()


def test_integer_background_form_is_exact():
    """The GPU walker evaluates the background-removed value as max(raw - ceil(S / 24), 0), S the
    integer sum of the three background sums; it must equal the reference's float expression
    trunc(max(raw - ((b0/8 + b1/8) + b2/8) / 3, 0)) for every S and raw (checked exhaustively over
    S at the raw values where the result changes, and on random triples)."""
    f32 = np.float32
    S = np.arange(0, 3 * 8 * 65535 + 1, dtype=np.int64)
    b0 = S // 3
    b1 = S // 3
    b2 = S - 2 * (S // 3)

    def float_form(raw, b0, b1, b2):
        s = (b0.astype(f32) / f32(8) + b1.astype(f32) / f32(8)) + b2.astype(f32) / f32(8)
        mb = s / f32(3)
        v = np.maximum(raw.astype(f32) - mb, f32(0))
        return v.astype(np.int64)

    for d in (-1, 0, 1, 2):
        raw = np.clip(S // 24 + d, 0, 65535)
        ref = float_form(raw, b0, b1, b2)
        got = np.maximum(raw - (b0 + b1 + b2 + 23) // 24, 0)
        assert np.array_equal(ref, got)
    rng = np.random.default_rng(5)
    bb = rng.integers(0, 8 * 65536, (3, 2_000_000))
    raw = rng.integers(0, 65536, 2_000_000)
    ref = float_form(raw, bb[0], bb[1], bb[2])
    got = np.maximum(raw - (bb.sum(0) + 23) // 24, 0)
    assert np.array_equal(ref, got)
