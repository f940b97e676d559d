"""The persistent work-queue fitter (cfq.hip) against the one-pass fitter (cfit_kernel).

Both run the same arithmetic per BFGS step, so parameters, iteration counts, flags and FP64
log-likelihoods must be bit-identical whatever order the queue hands the ROIs out in; the
evaluation count of the queue path is lower (an accepted trial's gradient is reused).  The
one-pass kernel itself is checked against the FP64 oracle in test_gpu_cfit.py."""
import numpy as np
import pytest
import torch

import qcstorm as Q

pytestmark = pytest.mark.gpu


def _sim(R, n, seed, kmax=3):
    sp = Q.default_sim(width=R, height=R, n_per_unit_max=kmax, min_sep=1.0, max_sep=4.0, photons_min=500.0,
                       photons_max=5000.0)
    rois, truth = Q.sim_rois_host(sp, seed, n)
    return rois, truth


def _chain(ctx, d_rois, n):
    """K = 1 (heuristic), K = 2 and K = 3 (previous fit plus an emitter at the residual maximum)."""
    out, prev = [], None
    for K in (1, 2, 3):
        P = 3 * K + 2
        th = torch.full((n, P), -7.0, dtype=torch.float32, device="cuda")
        info = torch.full((n, 4), -1, dtype=torch.int32, device="cuda")
        ll = torch.zeros(n, dtype=torch.float64, device="cuda")
        mode = Q.CFIT_INIT_HEUR if K == 1 else Q.CFIT_INIT_ADD
        ctx.cfit_dev(K, d_rois, n, mode, prev, th, info, ll)
        prev = th
        out.append((th, info, ll))
    torch.cuda.synchronize()
    return [(t.cpu().numpy(), i.cpu().numpy(), l.cpu().numpy()) for t, i, l in out]


@pytest.mark.parametrize("R,n", [(7, 1000), (9, 577), (11, 301), (7, 1)])
def test_queue_matches_one_pass_bitwise(gpu, R, n):
    rois, _ = _sim(R, n, 4100 + R)
    d_rois = torch.from_numpy(rois.view(np.int16)).cuda()
    ctx = Q.Context(Q.default_params(QE=1.0, ROISize=R))
    ctx.set_tuning(queue=0)
    ref = _chain(ctx, d_rois, n)
    for wpc in (0, 1):   # all resident waves / one wave per CU (many refills per lane)
        ctx.set_tuning(queue=1, waves_per_cu=wpc)
        got = _chain(ctx, d_rois, n)
        for K, ((t0, i0, l0), (t1, i1, l1)) in enumerate(zip(ref, got), start=1):
            print("R", R, "K", K, "wpc", wpc, "iters one-pass/queue", i0[:, 0].sum(), i1[:, 0].sum(),
                  "evals", i0[:, 2].sum(), i1[:, 2].sum())
            np.testing.assert_array_equal(t1.view(np.uint32), t0.view(np.uint32))
            np.testing.assert_array_equal(i1[:, :2], i0[:, :2])
            np.testing.assert_array_equal(l1.view(np.uint64), l0.view(np.uint64))
            assert (i1[:, 2] <= i0[:, 2]).all()


def test_queue_given_start_matches_one_pass(gpu):
    R, n = 9, 700
    rois, truth = _sim(R, n, 77)
    d_rois = torch.from_numpy(rois.view(np.int16)).cuda()
    # start: the simulated truth of the first two emitters, shifted
    init = np.zeros((n, 8), np.float32)
    for k in range(2):
        init[:, 3 * k] = np.maximum(truth[:, k, 4], 100.0) * 0.8
        init[:, 3 * k + 1] = np.where(truth[:, k, 4] > 0, truth[:, k, 1] + 0.3, R / 2)
        init[:, 3 * k + 2] = np.where(truth[:, k, 4] > 0, truth[:, k, 2] - 0.2, R / 2 + 1)
    init[:, 6], init[:, 7] = 1.4, 20.0
    d_init = torch.from_numpy(init).cuda()
    res = []
    for queue in (0, 1):
        ctx = Q.Context(Q.default_params(QE=1.0, ROISize=R))
        ctx.set_tuning(queue=queue)
        th = torch.zeros((n, 8), dtype=torch.float32, device="cuda")
        info = torch.zeros((n, 4), dtype=torch.int32, device="cuda")
        ll = torch.zeros(n, dtype=torch.float64, device="cuda")
        ctx.cfit_dev(2, d_rois, n, Q.CFIT_INIT_GIVEN, d_init, th, info, ll)
        torch.cuda.synchronize()
        res.append((th.cpu().numpy(), info.cpu().numpy(), ll.cpu().numpy()))
    np.testing.assert_array_equal(res[1][0].view(np.uint32), res[0][0].view(np.uint32))
    np.testing.assert_array_equal(res[1][1][:, :2], res[0][1][:, :2])
    np.testing.assert_array_equal(res[1][2].view(np.uint64), res[0][2].view(np.uint64))


def _fit_mle(R, rois, n, queue, kmax=3):
    ctx = Q.Context(Q.default_params(QE=1.0, ROISize=R, PixelSize=100.0))
    ctx.set_tuning(queue=queue)
    d_rois = torch.from_numpy(rois.view(np.int16)).cuda()
    loc = torch.zeros((3 * n, 12), dtype=torch.float32, device="cuda")
    sel = torch.zeros(n, dtype=torch.int32, device="cuda")
    ctx.fit_mle_dev(d_rois, n, kmax, loc, sel)
    torch.cuda.synchronize()
    return loc.cpu().numpy(), sel.cpu().numpy()


def test_mixed_sizes_match_per_size_calls(gpu):
    sizes = [(7, 1500), (11, 333), (9, 0), (9, 640)]   # an empty set is allowed
    data = {R: _sim(R, n, 5000 + R)[0] for R, n in sizes if n}
    for kmax in (3, 2):
        ctx = Q.Context(Q.default_params(QE=1.0, ROISize=7, PixelSize=100.0))
        sets, outs = [], []
        for R, n in sizes[1:] if kmax == 2 else [sizes[0], sizes[1], sizes[3]]:
            d_rois = torch.from_numpy(data[R].view(np.int16)).cuda() if n else torch.zeros(1, dtype=torch.int16, device="cuda")
            loc = torch.zeros((max(3 * n, 1), 12), dtype=torch.float32, device="cuda")
            sel = torch.zeros(max(n, 1), dtype=torch.int32, device="cuda")
            sets.append((R, d_rois, n, loc, sel))
            outs.append((R, n, loc, sel))
        ctx.fit_mle_mixed_dev(sets, kmax)
        torch.cuda.synchronize()
        for R, n, loc, sel in outs:
            if n == 0:
                continue
            rl, rs = _fit_mle(R, data[R], n, queue=1, kmax=kmax)
            print("kmax", kmax, "R", R, "n", n, "selected", np.bincount(rs, minlength=4))
            np.testing.assert_array_equal(sel.cpu().numpy()[:n], rs)
            np.testing.assert_array_equal(loc.cpu().numpy()[:3 * n].view(np.uint32), rl.view(np.uint32))


def test_fit_mle_queue_matches_one_pass(gpu):
    R, n = 7, 2000
    rois, _ = _sim(R, n, 6100)
    l0, s0 = _fit_mle(R, rois, n, queue=0)
    l1, s1 = _fit_mle(R, rois, n, queue=1)
    np.testing.assert_array_equal(s1, s0)
    np.testing.assert_array_equal(l1.view(np.uint32), l0.view(np.uint32))


def test_warm_path_queue_matches_one_pass(gpu):
    """The warm path (device-counted work lists, fallback list) gives the same rows either way."""
    n = 3000
    rois, _ = _sim(7, n, 6200)
    d_rois = torch.from_numpy(rois.view(np.int16)).cuda()
    net = Q.Net(seed=5)   # untrained: many low-confidence predictions take the fallback list
    res = []
    for queue in (0, 1):
        ctx = Q.Context(Q.default_params(QE=1.0, ROISize=7, PixelSize=100.0))
        ctx.set_tuning(queue=queue, warm_pmin=0.34)
        loc = torch.zeros((3 * n, 12), dtype=torch.float32, device="cuda")
        sel = torch.zeros(n, dtype=torch.int32, device="cuda")
        st = ctx.fit_warm_dev(net, d_rois, n, loc, sel)
        torch.cuda.synchronize()
        print("queue", queue, "stats", st)
        res.append((loc.cpu().numpy(), sel.cpu().numpy(), st))
    np.testing.assert_array_equal(res[1][1], res[0][1])
    np.testing.assert_array_equal(res[1][0].view(np.uint32), res[0][0].view(np.uint32))
    assert res[1][2]["fallbacks"] == res[0][2]["fallbacks"]


def test_tuning_rejects_bad_queue_settings(gpu):
    ctx = Q.Context(Q.default_params(QE=1.0, ROISize=7))
    with pytest.raises(RuntimeError):
        ctx.set_tuning(queue=2)
    with pytest.raises(RuntimeError):
        ctx.set_tuning(waves_per_cu=-1)
    t = ctx.tuning()
    assert t["queue"] == 1 and t["waves_per_cu"] == 0


def _valid(t, K, R, n_min=50.0):
    ok = np.ones(len(t), bool)
    for k in range(K):
        ok &= t[:, 3 * k] >= n_min
        ok &= (t[:, 3 * k + 1] > 0) & (t[:, 3 * k + 1] < R) & (t[:, 3 * k + 2] > 0) & (t[:, 3 * k + 2] < R)
    return ok & (t[:, 3 * K] > 0.45) & (t[:, 3 * K] < 0.45 * R)


@pytest.mark.parametrize("R", [7, 11])
def test_three_emitter_pruning_is_exact(gpu, R):
    """qcs_fit_mle_dev fits K = 3 only for ROIs that the first stage leaves at K = 2; the chosen
    counts must equal the full rule applied to K = 1, 2, 3 fits of every ROI."""
    n = 1500
    rois, _ = _sim(R, n, 7300 + R)
    d_rois = torch.from_numpy(rois.view(np.int16)).cuda()
    ctx = Q.Context(Q.default_params(QE=1.0, ROISize=R, PixelSize=100.0))
    (t1, _, l1), (t2, _, l2), (t3, _, l3) = _chain(ctx, d_rois, n)
    K = np.where(_valid(t1, 1, R), 1, 0)
    K = np.where(_valid(t2, 2, R) & ((K == 0) | (l2 - l1 > np.float64(np.float32(8.14)))), 2, K)
    K = np.where((K == 2) & _valid(t3, 3, R) & (l3 - l2 > np.float64(np.float32(8.14))), 3, K)
    loc = torch.zeros((3 * n, 12), dtype=torch.float32, device="cuda")
    sel = torch.zeros(n, dtype=torch.int32, device="cuda")
    st = ctx.fit_plain_stats_dev(d_rois, n, 3, loc, sel)
    got = sel.cpu().numpy()
    print("R", R, "counts", np.bincount(K, minlength=4), "stats", st)
    np.testing.assert_array_equal(got, K)
    assert st["final_k3"] == (K == 3).sum() and st["final_k2"] == (K == 2).sum()
