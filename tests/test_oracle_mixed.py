"""Pin the oracle's mixed-size extraction (oracle/orc_mixed.c): the size rule on hand-built mark
images whose answer follows from the rule's statement, and whole-frame properties that an
independent numpy restatement fixes (the candidates are those of the size-7 extraction; every
record is the frame crop at its size with the reference's header)."""
import numpy as np
import pytest

import oracle_lib as O
import qcstorm as Q


def _wle(s):
    return np.array([100.0, s, s + 0.3, s + 0.1, s + 0.2, 0.0], np.float32)   # min of the four = s


def _marks(w, h, pts):
    m = np.zeros((h, w), np.uint8)
    for x, y in pts:
        m[y, x] = 1
    return m


@pytest.mark.parametrize("s,R", [(1.0, 7), (1.4, 7), (1.41, 9), (1.9, 9), (1.91, 11), (3.0, 11), (float("nan"), 11)])
def test_size_from_psf_width(s, R):
    m = _marks(64, 64, [(30, 30)])
    assert O.mixed_size(m, 64, 30, 30, _wle(s)) == R


@pytest.mark.parametrize("d,R", [(1, 7), (2, 7), (3, 9), (4, 11), (5, 7)])
def test_size_from_nearest_mark(d, R):
    # a neighbour at Chebyshev distance d (diagonal or straight gives the same distance)
    for nb in ((30 + d, 30), (30 - d, 30 + d), (30, 30 - d)):
        m = _marks(64, 64, [(30, 30), nb])
        assert O.mixed_size(m, 64, 30, 30, _wle(1.2)) == R


def test_nearest_mark_wins_and_width_floor():
    m = _marks(64, 64, [(30, 30), (33, 30), (34, 34)])     # nearest at 3 -> 9 (not 11)
    assert O.mixed_size(m, 64, 30, 30, _wle(1.2)) == 9
    m = _marks(64, 64, [(30, 30), (34, 30)])               # 11 from the neighbour beats 9 from the width
    assert O.mixed_size(m, 64, 30, 30, _wle(1.6)) == 11
    m = _marks(64, 64, [(30, 30), (32, 30)])               # neighbour inside: the width decides
    assert O.mixed_size(m, 64, 30, 30, _wle(1.6)) == 9


def test_crowded_windows_shrink():
    # 3 other marks within the 11 x 11 window (two of them within 9 x 9): 11 -> 9
    m = _marks(64, 64, [(30, 30), (34, 30), (26, 26), (35, 35)])
    assert O.mixed_size(m, 64, 30, 30, _wle(1.2)) == 9
    # 3 other marks within the 9 x 9 window too: down to 7
    m = _marks(64, 64, [(30, 30), (34, 30), (26, 26), (33, 33)])
    assert O.mixed_size(m, 64, 30, 30, _wle(1.2)) == 7
    # marks in other frames of a stacked batch do not count (frame height 32: rows 32.. are frame 2)
    m = _marks(64, 64, [(30, 28), (30, 32), (30, 33)])
    assert O.mixed_size(m, 32, 30, 28, _wle(1.2)) == 7


@pytest.mark.parametrize("x,y,R", [(5, 30, 11), (4, 30, 9), (3, 30, 7), (30, 58, 11), (30, 59, 9), (60, 30, 7)])
def test_window_stays_in_the_frame(x, y, R):
    m = _marks(64, 64, [(x, y)])
    assert O.mixed_size(m, 64, x, y, _wle(2.5)) == R


def test_whole_frames_against_numpy():
    W, H, nf = 200, 160, 3
    sp = Q.default_sim(width=W, height=H, density_per_px=0.02, density_max_per_px=0.05, n_per_unit_max=3, min_sep=1.0,
                       max_sep=2.5)
    frames = Q.sim_frames_host(sp, 9, 1, nf)
    p = Q.default_params(QE=1.0, ImageWidth=W, ImageHigh=H, MultiEmitterFitEn=1, ROISize=7, BatchedImageNum=2)
    sets, sizes, nc, nfound = O.extract_mixed(p, frames, 11)
    r7, w7, found7, _ = O.extract(p, frames, 11)
    # the same candidates as the size-7 extraction, each in exactly one set
    key7 = [(int(r[49 + 2]) + 65536 * int(r[49 + 3]), int(r[49 + 1]), int(r[49])) for r in r7]
    keys = []
    for k, rois in enumerate(sets):
        R = 7 + 2 * k
        for r in rois:
            fr, yc, xc = int(r[R * R + 2]) + 65536 * int(r[R * R + 3]), int(r[R * R + 1]), int(r[R * R])
            keys.append((fr, yc, xc))
            crop = frames[fr - 11, yc - R // 2:yc + R // 2 + 1, xc - R // 2:xc + R // 2 + 1]
            np.testing.assert_array_equal(r[:R * R].reshape(R, R), crop)
            assert r[R * R + 4] in (0, 1) and not r[R * R + 5:].any()
        # raster order within each set
        ks = keys[len(keys) - rois.shape[0]:]
        assert ks == sorted(ks)
    assert sorted(keys) == sorted(key7) and nc == len(key7) == found7
    assert np.array_equal(np.bincount(sizes, minlength=12)[[7, 9, 11]], [s.shape[0] for s in sets])
    # the 7 x 7 set's records equal the size-7 extraction's records of the same candidates
    k7 = {k: i for i, k in enumerate(key7)}
    for r in sets[0]:
        i = k7[(int(r[51]) + 65536 * int(r[52]), int(r[50]), int(r[49]))]
        np.testing.assert_array_equal(r, r7[i])
