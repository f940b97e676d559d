"""Per-batch HIP graphs of the multi-emitter paths: every replay, with eager work and
synchronisations between replays, must equal the eager call on the same ROIs bit for bit (the
graph clears its counters with fill kernels, never with memset nodes)."""
import numpy as np
import pytest
import torch

import qcstorm as Q

pytestmark = pytest.mark.gpu


def _sim(n, seed):
    sp = Q.default_sim(width=7, height=7, n_per_unit_max=3, min_sep=1.0, max_sep=4.0, photons_min=500.0,
                       photons_max=5000.0)
    rois, _ = Q.sim_rois_host(sp, seed, n)
    return torch.from_numpy(rois.view(np.int16)).cuda()


@pytest.mark.parametrize("warm", [False, True])
def test_graph_replays_match_eager(gpu, warm):
    n = 3001
    p = Q.default_params(QE=1.0, ROISize=7, PixelSize=100.0)
    ctx = Q.Context(p)
    net = Q.Net(seed=4) if warm else None
    if warm:
        ctx.set_tuning(warm_pmin=0.5)
    buf = _sim(n, 100)
    loc = torch.zeros((3 * n, 12), dtype=torch.float32, device="cuda")
    sel = torch.zeros(n, dtype=torch.int32, device="cuda")
    g = ctx.graph_mle(buf, n, 3, loc, sel, net=net)
    for rep, seed in enumerate((101, 102, 101, 103)):
        new = _sim(n, seed)
        buf.copy_(new)
        torch.cuda.synchronize()
        g.launch()
        torch.cuda.synchronize()
        got_l, got_s = loc.cpu().numpy(), sel.cpu().numpy()
        # eager reference on the caller's context (other scratch), between replays
        ref_l = torch.zeros_like(loc)
        ref_s = torch.zeros_like(sel)
        if warm:
            ctx.fit_warm_dev(net, new, n, ref_l, ref_s, stats=False)
        else:
            ctx.fit_mle_dev(new, n, 3, ref_l, ref_s)
        torch.cuda.synchronize()
        print("replay", rep, "seed", seed, "counts", np.bincount(got_s, minlength=4))
        np.testing.assert_array_equal(got_s, ref_s.cpu().numpy())
        np.testing.assert_array_equal(got_l.view(np.uint32), ref_l.cpu().numpy().view(np.uint32))
    g.close()
