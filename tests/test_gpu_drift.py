"""GPU drift correction (qcs_drift_correct_dev) vs the oracle: the same frame count and groups,
group and per-frame shifts within 1e-4 px (the render sums with float atomics and the correlation
in a different order, so the correlation values differ in the last bits), corrected X/Y likewise,
the other fields untouched; at 2048 x 2048 SR pixels, recovery of a planted drift."""
import numpy as np
import pytest
import torch

import oracle_lib as O
import qcstorm as Q
from drift_data import drifting_rows

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("W,nf,corr,ax,wob", [(48, 60, 20, 0.01, 0.15), (40, 100, 20, 0.05, 0.0), (56, 45, 12, -0.02, 0.1)])
def test_drift_matches_oracle(gpu, W, nf, corr, ax, wob):
    rows, dx, dy, _ = drifting_rows(W, W, nf, 60, 3 + nf, ax=ax, wobble=wob)
    p = Q.default_params(ImageWidth=W, ImageHigh=W, PixelSize=100.0, SNR_th=5.0)
    ref, rfx, rfy, rtotal, rgx, rgy = O.drift_correct(p, rows, corr)
    ctx = Q.Context(p)
    d = torch.from_numpy(rows).cuda()
    total, fx, fy, gxy = ctx.drift_correct_dev(d, rows.shape[0], corr)
    got = d.cpu().numpy()
    print("case", W, nf, corr, ax, "groups", gxy.shape[0], "max |dg|", np.abs(gxy[:, 0] - rgx).max(), np.abs(gxy[:, 1] - rgy).max())
    assert total == rtotal == nf and gxy.shape[0] == rgx.shape[0]
    np.testing.assert_allclose(gxy[:, 0], rgx, atol=1e-4)
    np.testing.assert_allclose(gxy[:, 1], rgy, atol=1e-4)
    np.testing.assert_allclose(fx, rfx, atol=1e-4)
    np.testing.assert_allclose(fy, rfy, atol=1e-4)
    np.testing.assert_allclose(got[:, 1:3], ref[:, 1:3], atol=1e-4)
    keep = [0, 3, 4, 5, 6, 7, 8, 9, 10, 11]
    np.testing.assert_array_equal(got[:, keep], rows[:, keep])
    if ax == 0.05:   # 5 px of drift: later groups are found around the previous group's shift
        assert abs(gxy[-1, 0]) > 3.0


def test_drift_recovers_planted_drift_full_size(gpu):
    W = 256   # 2048 x 2048 SR pixels
    rows, dx, dy, _ = drifting_rows(W, W, 120, 3000, 9, ax=0.02, ay=-0.01, wobble=0.0)
    p = Q.default_params(ImageWidth=W, ImageHigh=W, PixelSize=100.0, SNR_th=5.0)
    ctx = Q.Context(p)
    d = torch.from_numpy(rows).cuda()
    total, fx, fy, gxy = ctx.drift_correct_dev(d, rows.shape[0], 30)
    gs, ge = O.drift_groups(120, 30)
    mdx = np.array([dx[s - 1:e].mean() for s, e in zip(gs, ge)])
    mdy = np.array([dy[s - 1:e].mean() for s, e in zip(gs, ge)])
    np.testing.assert_allclose(gxy[:, 0], mdx - mdx[0], atol=0.04)
    np.testing.assert_allclose(gxy[:, 1], mdy - mdy[0], atol=0.04)
    # after correction the rows sit at the structure (up to group 0's mean drift), not along the
    # drift: X - fx = (point + dx) - fx ~ point + mean drift of group 0
    got = d.cpu().numpy()
    f = rows[:, 11].astype(int)
    late = f > 60
    point_x = rows[late, 1] - dx[f[late] - 1]
    before = np.abs(rows[late, 1] - point_x - mdx[0]).mean()
    after = np.abs(got[late, 1] - point_x - mdx[0]).mean()
    assert after < 0.1 * before and after < 0.06


def test_drift_argument_errors(gpu):
    rows, _, _, _ = drifting_rows(40, 40, 15, 30, 5)
    p = Q.default_params(ImageWidth=40, ImageHigh=40)
    ctx = Q.Context(p)
    d = torch.from_numpy(rows).cuda()
    with pytest.raises(Q.QcsError):
        ctx.drift_correct_dev(d, rows.shape[0], 20)          # 15 frames < one group of 20
    with pytest.raises(Q.QcsError):
        ctx.drift_correct_dev(d, rows.shape[0], 10, shift_cap=5)
    np.testing.assert_array_equal(d.cpu().numpy(), rows)    # nothing changed
