"""GPU consecutive-frame merging (qcs_consec_fit_dev) vs the oracle's literal restatement of the
reference: the rows given back by every call equal the oracle's (bit for bit; NaN rows, which the
reference creates from zeroed rows, compared as NaN), so do the rows kept by the compaction after
it, the on-time ratios and the held-back counts, over sequences of calls."""
import numpy as np
import pytest
import torch

import oracle_lib as O
import qcstorm as Q

pytestmark = pytest.mark.gpu


def _blinking_rows(n_mol, nframes, seed, W=256.0, H=256.0, jitter=0.25, mean_on=2.5):
    """Molecules that stay on for a geometric number of frames, one row per frame on, jittered;
    rows sorted by frame, shuffled within each frame."""
    rng = np.random.default_rng(seed)
    rows = []
    for _ in range(n_mol):
        f0 = int(rng.integers(1, nframes + 1))
        d = int(rng.geometric(1.0 / mean_on))
        x, y = rng.uniform(2, W - 2), rng.uniform(2, H - 2)
        for f in range(f0, min(nframes, f0 + d - 1) + 1):
            r = np.zeros(12, np.float32)
            r[0] = rng.uniform(100, 3000)
            r[1], r[2] = x + rng.normal(0, jitter), y + rng.normal(0, jitter)
            r[3] = rng.uniform(-300, 300)
            r[4:6] = rng.uniform(1.0, 1.6, 2)
            r[6] = r[0] * rng.uniform(5, 8)
            r[7] = rng.uniform(10, 90)
            r[8] = rng.uniform(5, 40)
            r[9] = rng.uniform(4, 30)
            r[10] = rng.uniform(4, 30)
            r[11] = f
            rows.append(r)
    rows = np.stack(rows)
    key = rows[:, 11] + rng.uniform(0, 0.5, rows.shape[0])
    return rows[np.argsort(key, kind="stable")]


def _same(a, b):
    assert a.shape == b.shape
    nan = np.isnan(a)
    assert np.array_equal(nan, np.isnan(b))
    assert np.array_equal(a.view(np.uint32)[~nan], b.view(np.uint32)[~nan])


def _run(params, rows, cuts, ends=None):
    """Feed the frame groups [cuts[k], cuts[k+1]) to the GPU merger and the oracle; compare."""
    ctx = Q.Context(params)
    cs = Q.Consec(ctx)
    ref = O.Consec(params)
    fr = rows[:, 11]
    total_kept = 0
    ng = len(cuts) - 1
    for k in range(ng):
        g = rows[(fr >= cuts[k]) & (fr < cuts[k + 1])]
        is_end = (k == ng - 1) if ends is None else ends[k]
        want, wratio = ref.fit(g, is_end)
        d_in = torch.from_numpy(np.ascontiguousarray(g)).cuda() if g.shape[0] else torch.zeros((1, 12), device="cuda")
        d_out = torch.zeros((cs.held() + g.shape[0] + 1, 12), dtype=torch.float32, device="cuda")
        m, ratio = cs.fit_dev(d_in, g.shape[0], is_end, d_out)
        got = d_out[:m].cpu().numpy()
        _same(got, want)
        np.testing.assert_array_equal(ratio, wratio)
        assert cs.held() == ref.held()
        # removal of the zeroed rows (ZeroLocalizationsRemove, unsorted)
        kept = torch.zeros_like(d_out)
        km = ctx.compact_dev(d_out, m, kept)
        torch.cuda.synchronize()
        wk = O.compact(want)
        assert km == wk.shape[0]
        assert np.array_equal(kept[:km].cpu().numpy().view(np.uint32), wk.view(np.uint32))
        total_kept += km
    cs.close()
    return total_kept


@pytest.mark.parametrize("th_nm,lt", [(80.0, 0), (150.0, 0), (250.0, 1)])
def test_consec_matches_oracle(gpu, th_nm, lt):
    rows = _blinking_rows(4000, 40, 7 + int(th_nm), jitter=0.3)
    p = Q.default_params(ImageWidth=256, ImageHigh=256, ConsecFitEn=1, ConsecFit_DistanceTh_nm=th_nm, PixelSize=100.0,
                         LocType=lt, QE=0.77)
    kept = _run(p, rows, [1, 6, 13, 14, 25, 41])
    assert 0 < kept < rows.shape[0]


def test_consec_dense_ties_and_shared_tails(gpu):
    """A crowded field on a coarse grid: many equal integer distances and rows linked by two."""
    rng = np.random.default_rng(3)
    n = 20000
    rows = np.zeros((n, 12), np.float32)
    rows[:, 11] = np.sort(rng.integers(1, 21, n))
    rows[:, 1] = rng.integers(10, 40, n) + rng.uniform(0, 1, n)
    rows[:, 2] = rng.integers(10, 40, n) + rng.uniform(0, 1, n)
    rows[:, 0] = rng.uniform(100, 1000, n)
    rows[:, 4:6] = 1.3
    rows[:, 6] = rows[:, 0] * 6
    rows[:, 7] = 30
    rows[:, 9] = rng.uniform(5, 20, n)
    rows[:, 10] = rows[:, 9]
    p = Q.default_params(ImageWidth=64, ImageHigh=64, ConsecFitEn=1, ConsecFit_DistanceTh_nm=180.0, PixelSize=100.0)
    _run(p, rows, [1, 8, 21])


def test_consec_single_call_and_empty_calls(gpu):
    rows = _blinking_rows(500, 10, 11)
    p = Q.default_params(ImageWidth=256, ImageHigh=256, ConsecFitEn=1, ConsecFit_DistanceTh_nm=100.0)
    _run(p, rows, [1, 11])                                              # one end call
    _run(p, rows, [1, 4, 4, 11, 11], ends=[False, False, False, True])  # empty groups in between
    _run(p, rows, [1, 5, 11, 11], ends=[False, False, True])            # the end call gives back only held rows


def test_consec_large_batch(gpu):
    rows = _blinking_rows(60000, 60, 5, W=2048.0, H=2048.0)
    p = Q.default_params(ImageWidth=2048, ImageHigh=2048, ConsecFitEn=1, ConsecFit_DistanceTh_nm=100.0)
    _run(p, rows, [1, 31, 61])


def test_consec_rejects_unsorted_rows_and_keeps_state(gpu):
    rows = _blinking_rows(800, 12, 13)
    p = Q.default_params(ImageWidth=256, ImageHigh=256, ConsecFitEn=1, ConsecFit_DistanceTh_nm=100.0)
    ctx = Q.Context(p)
    cs = Q.Consec(ctx)
    ref = O.Consec(p)
    g1, g2 = rows[rows[:, 11] < 6], rows[rows[:, 11] >= 6]
    d_out = torch.zeros((rows.shape[0] + 1, 12), dtype=torch.float32, device="cuda")
    cs.fit_dev(torch.from_numpy(g1).cuda(), g1.shape[0], False, d_out)
    ref.fit(g1, False)
    bad = g2[::-1].copy()
    with pytest.raises(Q.QcsError):
        cs.fit_dev(torch.from_numpy(bad).cuda(), bad.shape[0], False, d_out)
    neg = g2.copy()
    neg[0, 11] = -3
    with pytest.raises(Q.QcsError):
        cs.fit_dev(torch.from_numpy(neg).cuda(), neg.shape[0], False, d_out)
    assert cs.held() == g1.shape[0]
    m, _ = cs.fit_dev(torch.from_numpy(g2).cuda(), g2.shape[0], True, d_out)
    want, _ = ref.fit(g2, True)
    _same(d_out[:m].cpu().numpy(), want)
    small = torch.zeros((1, 12), dtype=torch.float32, device="cuda")
    cs.fit_dev(torch.from_numpy(g1).cuda(), g1.shape[0], False, small)   # nothing given back yet
    with pytest.raises(Q.QcsError):
        cs.fit_dev(torch.from_numpy(g2).cuda(), g2.shape[0], False, small)   # out_cap < held rows
