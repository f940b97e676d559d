"""The autotuner of the multi-emitter paths: it must pick the fastest candidate within the
tolerances, install it as the context's tuning, and the near-reference candidates must agree with
the converged reference."""
import numpy as np
import pytest
import torch

import qcstorm as Q

pytestmark = pytest.mark.gpu


def test_autotune_picks_fastest_within_tolerance(gpu):
    n = 1 << 15
    sp = Q.default_sim(width=7, height=7, n_per_unit_max=3, min_sep=1.0, max_sep=4.0, photons_min=500.0,
                       photons_max=5000.0)
    rois, _ = Q.sim_rois_host(sp, 808, n)
    d_rois = torch.from_numpy(rois.view(np.int16)).cuda()
    ctx = Q.Context(Q.default_params(QE=1.0, ROISize=7, PixelSize=100.0))
    net = Q.Net(seed=2)
    tab = ctx.autotune_dev(d_rois, n, net=net, trials=1, min_select_agree=0.99, max_dx_p99=0.02)
    for e in tab:
        print(e)
    plain = [e for e in tab if not e["warm"]]
    assert len(plain) >= 12 and any(e["warm"] for e in tab)
    # a candidate at the reference caps without the loss-decrease stop reproduces the reference
    ref_like = [e for e in plain if e["max_iter"] == [100, 60, 60] and e["ftol"] == 0.0][0]
    assert ref_like["select_agree"] > 0.995 and ref_like["dx_p99"] < 0.01
    within = [e for e in plain if e["within"]]
    chosen = [e for e in plain if e["chosen"]]
    assert len(chosen) == 1 and chosen[0]["within"]
    assert chosen[0]["ms"] == min(e["ms"] for e in within)
    t = ctx.tuning()
    assert t["max_iter"] == chosen[0]["max_iter"] and abs(t["ftol"] - chosen[0]["ftol"]) < 1e-9
    assert t["queue"] == chosen[0]["queue"] and t["waves_per_cu"] == chosen[0]["waves_per_cu"]
