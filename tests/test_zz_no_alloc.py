"""The allocation-free fit entry point (qcs_fit_dev_noalloc) against qcs_fit_dev.

Both run the same kernels; the variant takes the multi-emitter cascade's scratch from a fixed
device array and leaves the row count on the device.  Its rows must equal qcs_fit_dev's bit for
bit, on two calls in a row with the second smaller (the array keeps the first call's contents,
so anything the cascade must start from zero has to be cleared by the call itself)."""
import numpy as np
import pytest
import torch

import oracle_lib as O
import qcstorm as Q

pytestmark = pytest.mark.gpu


def _multi_rois(R, n, seed):
    sp = Q.default_sim(width=R, height=R, n_per_unit_max=3, min_sep=1.0, max_sep=3.0, photons_min=800,
                       photons_max=6000)
    rois, _ = Q.sim_rois_host(sp, seed, n)
    rng = np.random.default_rng(seed)
    rois[:, R * R + 4] = (rng.random(n) < 0.3).astype(np.uint16)
    return rois


def _both(ctx, rois, wle, n):
    """(rows of qcs_fit_dev, rows of qcs_fit_dev_noalloc) for the first n ROIs."""
    d_rois = torch.from_numpy(rois[:n].view(np.int16)).cuda()
    d_wle = torch.from_numpy(wle[:n]).cuda()
    cap = 2 * n
    d_ref = torch.full((cap, 12), 7.0, dtype=torch.float32, device="cuda")
    d_got = torch.full((cap, 12), 7.0, dtype=torch.float32, device="cuda")
    d_cnt = torch.full((1,), -1, dtype=torch.int32, device="cuda")
    rows = ctx.fit_dev(d_rois, d_wle, n, d_ref)
    ctx.fit_dev_noalloc(d_rois, d_wle, n, d_got, d_cnt)
    torch.cuda.synchronize()
    got_rows = int(d_cnt.cpu().item())
    return rows, d_ref[:rows].cpu().numpy(), got_rows, d_got[:max(got_rows, 0)].cpu().numpy()


@pytest.mark.parametrize("R,lt", [(7, 0), (9, 0), (7, 1)])
def test_noalloc_matches_fit_dev_multi_emitter(gpu, R, lt):
    params = Q.default_params(QE=1.0, ROISize=R, LocType=lt, MultiEmitterFitEn=1)
    ctx = Q.Context(params)
    rois = _multi_rois(R, 5000, 900 + R + 10 * lt)
    wle = O.wle(rois, R, lt)
    for n in (5000, 1733):   # second call smaller: stale scratch from the first must not leak in
        rows, ref, got_rows, got = _both(ctx, rois, wle, n)
        st = ctx.stats()
        print("R", R, "lt", lt, "n", n, "rows", rows, got_rows, "2E", st["n_2e"], "3E", st["n_3e"])
        assert got_rows == rows and rows > n   # second emitters were appended
        np.testing.assert_array_equal(got.view(np.uint32), ref.view(np.uint32))
    ctx.close()


def test_noalloc_matches_fit_dev_single_emitter(gpu):
    params = Q.default_params(QE=1.0)
    ctx = Q.Context(params)
    sp = Q.default_sim()
    rois, _ = Q.sim_rois_host(sp, 1000, 4096)
    wle = O.wle(rois, 7)
    for n in (4096, 999):
        rows, ref, got_rows, got = _both(ctx, rois, wle, n)
        assert rows == n and got_rows == n
        np.testing.assert_array_equal(got.view(np.uint32), ref.view(np.uint32))
    ctx.close()


def test_noalloc_rejects_oversized_calls(gpu):
    ctx = Q.Context(Q.default_params(QE=1.0, MultiEmitterFitEn=1))
    n = Q.NOALLOC_MAX_ROIS + 1
    d = torch.zeros(1, dtype=torch.int32, device="cuda")
    loc = torch.zeros((1, 12), dtype=torch.float32, device="cuda")
    # rejected on the host before any launch (the buffers are never read)
    with pytest.raises(Q.QcsError) as e:
        ctx.fit_dev_noalloc(d, d, n, loc, d)
    assert e.value.rc == Q.QCS_ERR_CAPACITY
    ctx.close()
