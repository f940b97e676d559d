"""Pinning the oracle's consecutive-frame merging (oracle/orc_consec.c).

The link search is checked against an independent brute force (nearest row of the next frame, all
rows of that frame, on truncated positions) on frame-sorted rows, and against hand-built cases for
the threshold, ties, truncation and the 32-row block start; the merge against the closed-form
weighted average; the held-back group, the on-time ratios, the CRLB of the output rows and the
zero-row quirk against their definitions."""
import numpy as np
import pytest

import oracle_lib as O
import qcstorm as Q


def _row(x, y, f, pp=500.0, crb=10.0, tp=3000.0, bg=20.0, sx=1.3, sy=1.2, z=0.0):
    r = np.zeros(12, np.float32)
    r[0], r[1], r[2], r[3], r[4], r[5], r[6], r[7], r[8], r[9], r[10], r[11] = \
        pp, x, y, z, sx, sy, tp, bg, 20.0, crb, crb, f
    return r


def _sorted_rows(n, seed, nframes, W=64.0):
    rng = np.random.default_rng(seed)
    loc = np.zeros((n, 12), np.float32)
    loc[:, 11] = np.sort(rng.integers(1, nframes + 1, n))
    loc[:, 1] = rng.uniform(0, W, n)
    loc[:, 2] = rng.uniform(0, W, n)
    loc[:, 0] = rng.uniform(100, 2000, n)
    loc[:, 4:6] = rng.uniform(1.0, 1.6, (n, 2))
    loc[:, 6] = loc[:, 0] * 6
    loc[:, 7] = rng.uniform(10, 90, n)
    loc[:, 8] = rng.uniform(5, 40, n)
    loc[:, 9] = rng.uniform(4, 30, n)
    loc[:, 10] = rng.uniform(4, 30, n)
    return loc


def _brute_links(loc, th):
    xi = loc[:, 1].astype(np.int64)
    yi = loc[:, 2].astype(np.int64)
    fr = loc[:, 11].astype(np.int64)
    th2 = np.float32(th) * np.float32(th)
    n = loc.shape[0]
    back = np.full(n, -2, np.int64)
    fwd = np.zeros(n, np.int64)
    for i in range(n):
        cand = np.nonzero(fr == fr[i] + 1)[0]
        if cand.size == 0:
            continue
        d = ((xi[cand] - xi[i]) ** 2 + (yi[cand] - yi[i]) ** 2).astype(np.float32)
        ok = d < th2
        if not ok.any():
            continue
        c, d = cand[ok], d[ok]
        j = c[np.lexsort((c, d))[0]]   # smallest distance, then smallest index
        back[i] = j
        fwd[j] = 1
    return back, fwd


@pytest.mark.parametrize("th", [0.8, 1.5, 3.0])
def test_links_match_brute_force(th):
    loc = _sorted_rows(3000, 1, 25, W=40.0)   # dense: many candidates, many ties on integer positions
    back, fwd = O.consec_links(loc, th)
    bb, bf = _brute_links(loc, th)
    np.testing.assert_array_equal(back, bb)
    np.testing.assert_array_equal(fwd, bf)
    assert (back >= 0).sum() > 100


def test_link_threshold_ties_truncation():
    rows = np.stack([
        _row(10.9, 10.0, 1),    # 0: int (10, 10)
        _row(12.0, 10.0, 2),    # 1: int distance 2 -> not < 2
        _row(9.0, 10.0, 2),     # 2: int distance 1
        _row(11.99, 10.0, 2),   # 3: int (11, 10): distance 1, tie with row 2 -> row 2 (first) wins
        _row(10.0, 10.0, 4),    # 4: frame gap 2 from row 2/3
    ])
    back, fwd = O.consec_links(rows, 2.0)
    assert back.tolist() == [2, -2, -2, -2, -2]
    assert fwd.tolist() == [0, 0, 1, 0, 0]
    back, _ = O.consec_links(rows, 2.01)   # 2^2 < 2.01^2: row 1 qualifies but row 2 is closer
    assert back[0] == 2


def test_link_scan_starts_at_block():
    """A row scans from the first row of its 32-row block until it has read a row whose frame
    exceeds its own + 1.  With frame-sorted rows and zeroed rows (frame 0, position 0) among them,
    a zeroed row links only to frame-1 rows at or after its block's first row."""
    rows = np.stack([_row(20.0, 20.0, 1)] * 40)
    rows[2] = _row(0.5, 0.5, 1)      # frame 1 at int (0, 0), block 0
    rows[33] = 0.0                   # zeroed rows in block 1
    rows[34] = 0.0
    rows[35] = _row(0.2, 0.7, 1)     # frame 1 at int (0, 0), block 1
    rows[36:] = _row(20.0, 20.0, 2)
    back, fwd = O.consec_links(rows, 1.0)
    assert back[33] == 35 and back[34] == 35 and fwd[35] == 1 and fwd[2] == 0
    # out of frame order the scan stops early: a frame-5 row at the block start ends the scan of
    # a frame-0 row before it reaches the frame-1 row
    rows[32] = _row(20.0, 20.0, 5)
    back, _ = O.consec_links(rows, 1.0)
    assert back[33] == -2 and back[34] == -2


def test_merge_closed_form_and_ontime():
    p = Q.default_params(ConsecFit_DistanceTh_nm=150.0, PixelSize=100.0, QE=0.8)
    xs, crb = [10.2, 10.6, 10.4], [8.0, 12.0, 20.0]
    chain = [_row(x, 30.3, f, pp=400 + 10 * f, crb=c, tp=2000 + f, bg=15 + f, z=5.0 * f)
             for x, c, f in zip(xs, crb, (1, 2, 3))]
    single = _row(40.0, 40.0, 2, crb=9.0)
    rows = np.stack([chain[0], chain[1], single, chain[2]])
    cs = O.Consec(p)
    out, ratio = cs.fit(rows, is_end=True)
    assert out.shape == (4, 12)
    w = np.array([10.0 / c for c in crb])
    np.testing.assert_allclose(out[0, 1], np.dot(w, xs) / w.sum(), rtol=1e-6)
    np.testing.assert_allclose(out[0, 3], np.dot(w, [5.0, 10.0, 15.0]) / w.sum(), rtol=1e-6)
    assert out[0, 2] == pytest.approx(30.3, rel=1e-6)
    assert out[0, 0] == 1260 and out[0, 6] == 6006 and out[0, 7] == 51
    assert out[0, 8] == pytest.approx(np.sqrt(0.8) * 1260 / np.sqrt(1260 + 51), rel=1e-6)
    assert out[0, 11] == 1                          # frame of the head
    assert not out[1].any() and not out[3].any()    # chain members zeroed
    assert out[2, 1] == pytest.approx(40.0, rel=1e-6) and out[2, 0] == 500
    # on-time over the call's new rows: one chain of 3 frames (bin 2) and one single (bin 0)
    assert ratio[2] == 0.5 and ratio[0] == 0.5 and ratio.sum() == 1.0
    # the CRLB is recomputed only for rows held back from an earlier call (none here)
    np.testing.assert_array_equal(out[[0, 2], 9], [crb[0], 9.0])


def test_held_back_group_and_crlb():
    p = Q.default_params(ConsecFit_DistanceTh_nm=100.0, PixelSize=100.0)
    g1 = np.stack([_row(5.0, 5.0, 1), _row(30.0, 30.0, 2)])
    g2 = np.stack([_row(30.2, 30.4, 3), _row(50.0, 50.0, 4)])
    g3 = np.stack([_row(50.5, 50.5, 5), _row(10.0, 10.0, 5)])
    cs = O.Consec(p)
    out1, _ = cs.fit(g1, is_end=False)
    assert out1.shape[0] == 0 and cs.held() == 2
    out2, r2 = cs.fit(g2, is_end=False)     # gives back g1; (30,30)@2 absorbs (30.2,30.4)@3
    assert out2.shape[0] == 2 and cs.held() == 2
    assert out2[1, 0] == 1000 and out2[1, 1] == pytest.approx((30.0 + 30.2) / 2, rel=1e-6)
    # the rows given back carry the CRLB recomputed from their (merged) photons and widths
    np.testing.assert_array_equal(out2[:, 9:11], O.crlb(p, out2)[:, 9:11])
    assert out2[1, 9] < 10.0                # more photons: better than the input 10 nm
    out3, _ = cs.fit(g3, is_end=True)       # gives back g2 (its first row was zeroed) and g3
    assert out3.shape[0] == 4 and cs.held() == 0
    # the zeroed row comes back as a row of zero photons (NaN positions): compaction drops it
    assert out3[0, 0] == 0 and np.isnan(out3[0, 1])
    kept = O.compact(out3)
    assert kept.shape[0] == 2 and kept[0, 0] == 1000 and kept[0, 11] == 4 and kept[1, 11] == 5
    # rows held back from an earlier call get the recomputed CRLB; the end call's new rows keep
    # their input CRLB
    np.testing.assert_array_equal(out3[:2, 9], O.crlb(p, out3[:2])[:, 9])
    assert out3[3, 9] == 10.0 and O.crlb(p, out3[3:])[0, 9] != 10.0


def test_shared_tail_is_summed_into_both_heads():
    p = Q.default_params(ConsecFit_DistanceTh_nm=250.0, PixelSize=100.0)
    rows = np.stack([_row(10.0, 10.0, 1), _row(12.0, 10.0, 1), _row(11.0, 10.0, 2)])
    out, _ = O.Consec(p).fit(rows, is_end=True)
    assert out[0, 0] == 1000 and out[1, 0] == 1000 and not out[2].any()
