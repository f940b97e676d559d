"""GPU WLE pre-estimation (PSF widths, nearest neighbour, type) against the oracle: bit-exact."""
import numpy as np
import pytest
import torch

import oracle_lib as O
import qcstorm as Q

pytestmark = pytest.mark.gpu


def _rois_with_headers(R, n, seed, frames=7, shuffle=False):
    sp = Q.default_sim(width=R, height=R)
    rois, _ = Q.sim_rois_host(sp, seed, n)
    rng = np.random.default_rng(seed)
    # spread the ROIs over a few frames at random positions so the neighbour search has work
    f = np.sort(rng.integers(1, frames + 1, n))
    y = rng.integers(R, 500, n)
    x = rng.integers(R, 500, n)
    order = np.lexsort((x, y, f))
    P = R * R
    rois[:, P + 0] = x[order]
    rois[:, P + 1] = y[order]
    rois[:, P + 2] = f[order] % 65536
    rois[:, P + 3] = f[order] // 65536
    if shuffle:
        rois = rois[rng.permutation(n)]
    return rois


@pytest.mark.parametrize("R,shuffle", [(7, False), (7, True), (9, False), (11, False), (5, True)])
def test_wle_bit_exact(gpu, R, shuffle):
    rois = _rois_with_headers(R, 3000, 17 + R, shuffle=shuffle)
    ref = O.wle(rois, R)
    ctx = Q.Context(Q.default_params(ROISize=R))
    d_rois = torch.from_numpy(rois.view(np.int16)).cuda()
    d_wle = torch.zeros((rois.shape[0], Q.WLE_FLOATS), dtype=torch.float32, device="cuda")
    ctx.wle_dev(d_rois, rois.shape[0], d_wle)
    torch.cuda.synchronize()
    got = d_wle.cpu().numpy()
    bad = np.argwhere(got.view(np.uint32) != ref.view(np.uint32))
    print("R", R, "shuffle", shuffle, "types", np.bincount(ref[:, 5].astype(int), minlength=3), "mismatch", len(bad))
    if len(bad):
        i = bad[0][0]
        print("row", i, "gpu", got[i], "orc", ref[i])
    assert len(bad) == 0
