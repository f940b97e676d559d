"""Converged fitting of the elliptical (astigmatic 3D) model on the GPU vs the FP64 CPU oracle.

Model per emitter: N / (2 pi sx sy) exp(-(dx^2 / (2 sx^2) + dy^2 / (2 sy^2))), each emitter with its
own widths, plus a shared background: 5K + 1 parameters (16 at three emitters).  The GPU fitter
(BFGS, FP32) and the oracle's FP64 Levenberg-Marquardt MLE (oracle/orc_truth.c, astig = 1)
maximise the same Poisson likelihood, so started from the GPU result the oracle must stay within
the FP32 convergence tolerance.  The rows of qcs_fit_mle_dev carry the oracle's FP64 full-Fisher
CRLB at the same parameters and z from the calibration polynomial of sx^2 - sy^2; the persistent
queue gives the one-pass kernel's results bit for bit."""
import numpy as np
import pytest
import torch

import oracle_lib as O
import qcstorm as Q

pytestmark = pytest.mark.gpu

# sigma_x/y = 1.2 sqrt(1 + ((z -+ 250) / 400)^2) gives sx^2 - sy^2 = -0.009 z exactly
ZCAL = dict(p1_XGY=-111.11, p1_XLY=-111.11, ZDepthCorrFactor=1.0)


def _params(R, **kw):
    return Q.default_params(QE=1.0, ROISize=R, LocType=Q.LOCTYPE_AS3D, PixelSize=100.0, **ZCAL, **kw)


def _sim(R, n, seed, kmax=1, **kw):
    sp = Q.default_sim(width=R, height=R, astig=1, n_per_unit_max=kmax, **kw)
    return Q.sim_rois_host(sp, seed, n)


def _q(rois, R, p):
    img = rois[:, :R * R].astype(np.float64).reshape(-1, R, R)
    return np.maximum((img - p.Offset) * np.float32(p.KAdc), 1.0) / p.QE


def _valid(t, K, R):
    """The selection's validity rule: every emitter >= 50 photons inside the ROI, widths in (0.45, 0.45 R)."""
    for k in range(K):
        N, x, y, sx, sy = t[5 * k:5 * k + 5]
        if not (N >= 50 and 0 < x < R and 0 < y < R and 0.45 < sx < 0.45 * R and 0.45 < sy < 0.45 * R):
            return False
    return True


def _cfit(ctx, K, d_rois, n, mode=Q.CFIT_INIT_HEUR, init=None, fill=0.0):
    P = 5 * K + 1
    th = torch.full((n, P), fill, dtype=torch.float32, device="cuda")
    info = torch.full((n, 4), -1, dtype=torch.int32, device="cuda")
    ll = torch.zeros(n, dtype=torch.float64, device="cuda")
    ctx.cfit_dev(K, d_rois, n, mode, init, th, info, ll)
    return th, info, ll


@pytest.mark.parametrize("R", [7, 9, 11])
def test_elliptical_single_emitter_matches_fp64_oracle(gpu, R):
    n = 512
    rois, truth = _sim(R, n, 1300 + R, photons_min=1000.0, photons_max=6000.0, z_min=-400.0, z_max=400.0)
    p = _params(R)
    ctx = Q.Context(p)
    d_rois = torch.from_numpy(rois.view(np.int16)).cuda()
    th, info, ll = _cfit(ctx, 1, d_rois, n)
    torch.cuda.synchronize()
    th, info, ll = th.cpu().numpy(), info.cpu().numpy(), ll.cpu().numpy()
    conv = (info[:, 1] & Q.CFIT_CONVERGED) != 0
    print("R", R, "converged", conv.mean(), "iters p50/p95", np.percentile(info[:, 0], [50, 95]))
    assert conv.mean() > 0.95
    q = _q(rois, R, p)
    dx, ds, dn, dll = [], [], [], []
    for i in np.nonzero(conv)[0][:200]:
        t, it, l64 = O.mle(q[i], K=1, astig=True, theta0=th[i].astype(np.float64))
        dx.append(max(abs(t[1] - th[i, 1]), abs(t[2] - th[i, 2])))
        ds.append(max(abs(t[3] - th[i, 3]), abs(t[4] - th[i, 4])))
        dn.append(abs(t[0] - th[i, 0]) / t[0])
        dll.append(abs(l64 - ll[i]))
    dx, ds, dn, dll = map(np.array, (dx, ds, dn, dll))
    print("|dx| p99/max", np.percentile(dx, 99), dx.max(), "|dsigma| p99", np.percentile(ds, 99), "photons rel p99",
          np.percentile(dn, 99), "loglik diff max", dll.max())
    assert np.percentile(dx, 99) < 2e-3 and dx.max() < 1e-2
    assert np.percentile(ds, 99) < 2e-3
    assert np.percentile(dn, 99) < 2e-3
    assert dll.max() < 1e-2
    # the widths track the simulated ones and the positions the simulated emitters
    inner = conv & (np.abs(truth[:, 0, 3]) < 350)
    assert np.median(np.abs(th[inner, 3] - truth[inner, 0, 6])) < 0.08
    assert np.median(np.abs(th[inner, 4] - truth[inner, 0, 7])) < 0.08
    err = np.hypot(th[conv, 1] - truth[conv, 0, 1], th[conv, 2] - truth[conv, 0, 2])
    assert np.median(err) < 0.12


def test_elliptical_two_emitters_from_residual_init_match_fp64_oracle(gpu):
    R, n = 11, 384
    rois, truth = _sim(R, n, 1401, kmax=2, min_sep=3.0, max_sep=4.5, photons_min=2000.0, photons_max=6000.0,
                       z_min=-300.0, z_max=300.0)
    p = _params(R)
    ctx = Q.Context(p)
    d_rois = torch.from_numpy(rois.view(np.int16)).cuda()
    t1, _, _ = _cfit(ctx, 1, d_rois, n)
    t2, info, ll = _cfit(ctx, 2, d_rois, n, Q.CFIT_INIT_ADD, t1)
    torch.cuda.synchronize()
    t2, info = t2.cpu().numpy(), info.cpu().numpy()
    two = truth[:, 1, 4] > 0
    conv = ((info[:, 1] & Q.CFIT_CONVERGED) != 0) & two
    # fits that the selection would reject (an emitter on the ROI border or faded out) are not
    # optima the oracle is asked to confirm; they must stay rare
    val = conv & np.array([_valid(t2[i], 2, R) for i in range(n)])
    print("two-emitter ROIs", two.sum(), "converged", conv.sum(), "valid", val.sum())
    assert conv.sum() > 0.9 * two.sum() and val.sum() > 0.9 * two.sum()
    q = _q(rois, R, p)
    dx = []
    for i in np.nonzero(val)[0][:120]:
        t, it, l64 = O.mle(q[i], K=2, astig=True, theta0=t2[i].astype(np.float64))
        dx.append(np.abs(t[[1, 2, 6, 7]] - t2[i, [1, 2, 6, 7]]).max())
    dx = np.array(dx)
    print("2E max |dx| p50/p95/max", np.percentile(dx, [50, 95]), dx.max())
    assert np.percentile(dx, 95) < 1e-3 and dx.max() < 2e-2


def test_elliptical_fit_mle_rows_crlb_and_z(gpu):
    R, n = 11, 768
    rois, truth = _sim(R, n, 1502, kmax=3, min_sep=3.0, max_sep=4.5, photons_min=2500.0, photons_max=7000.0,
                       bg_min=10.0, bg_max=40.0, z_min=-300.0, z_max=300.0)
    p = _params(R)
    ctx = Q.Context(p)
    d_rois = torch.from_numpy(rois.view(np.int16)).cuda()
    loc = torch.zeros((3 * n, 12), dtype=torch.float32, device="cuda")
    sel = torch.zeros(n, dtype=torch.int32, device="cuda")
    ctx.fit_mle_dev(d_rois, n, 3, loc, sel)
    torch.cuda.synchronize()
    loc, sel = loc.cpu().numpy().reshape(n, 3, 12), sel.cpu().numpy()
    k_true = (truth[:, :, 4] > 0).sum(1)
    acc = {k: float((sel[k_true == k] == k).mean()) for k in (1, 2, 3)}
    print("selection accuracy by true count", acc, "counts", np.bincount(k_true))
    assert acc[1] > 0.95 and acc[2] > 0.85 and acc[3] > 0.65
    checked, dz = 0, []
    for i in range(n):
        K = sel[i]
        if K == 0:
            continue
        rows = loc[i, :K]
        th = []
        for k in range(K):
            th += [rows[k, 6], rows[k, 1] - 100 + R // 2, rows[k, 2] - 100 + R // 2, rows[k, 4], rows[k, 5]]
        th += [rows[0, 7]]
        cr = O.fisher_crlb(R, np.array(th, np.float64), K=K, astig=True)
        for k in range(K):
            sx, sy = float(rows[k, 4]), float(rows[k, 5])
            np.testing.assert_allclose(rows[k, 9], cr[5 * k + 1] * 100.0, rtol=5e-5)
            np.testing.assert_allclose(rows[k, 10], cr[5 * k + 2] * 100.0, rtol=5e-5)
            np.testing.assert_allclose(rows[k, 0], rows[k, 6] / (2 * np.pi * sx * sy), rtol=1e-5)
            np.testing.assert_allclose(rows[k, 3], -111.11 * (sx * sx - sy * sy), rtol=1e-4, atol=1e-3)
        if K == k_true[i] == 1:
            dz.append(rows[0, 3] - truth[i, 0, 3])
        checked += 1
        if checked >= 150:
            break
    assert checked >= 150
    dz = np.array(dz)
    print("single-emitter z error median |dz| nm", np.median(np.abs(dz)), "n", dz.size)
    assert dz.size > 20 and np.median(np.abs(dz)) < 40.0


@pytest.mark.parametrize("R,n", [(9, 577), (11, 301)])
def test_elliptical_queue_matches_one_pass_bitwise(gpu, R, n):
    rois, _ = _sim(R, n, 1600 + R, kmax=3, min_sep=1.5, max_sep=4.0, photons_min=800.0, photons_max=6000.0)
    d_rois = torch.from_numpy(rois.view(np.int16)).cuda()
    ctx = Q.Context(_params(R))

    def chain():
        out, prev = [], None
        for K in (1, 2, 3):
            th, info, ll = _cfit(ctx, K, d_rois, n, Q.CFIT_INIT_HEUR if K == 1 else Q.CFIT_INIT_ADD, prev, fill=-7.0)
            prev = th
            out.append((th, info, ll))
        torch.cuda.synchronize()
        return [(t.cpu().numpy(), i.cpu().numpy(), l.cpu().numpy()) for t, i, l in out]

    ctx.set_tuning(queue=0)
    ref = chain()
    ctx.set_tuning(queue=1)
    got = chain()
    for K, ((t0, i0, l0), (t1, i1, l1)) in enumerate(zip(ref, got), start=1):
        print("R", R, "K", K, "iters", i0[:, 0].sum(), "converged", ((i0[:, 1] & Q.CFIT_CONVERGED) != 0).mean())
        np.testing.assert_array_equal(t1.view(np.uint32), t0.view(np.uint32))
        np.testing.assert_array_equal(i1[:, :2], i0[:, :2])
        np.testing.assert_array_equal(l1.view(np.uint64), l0.view(np.uint64))


def test_elliptical_mixed_sizes_match_per_size_calls(gpu):
    sizes = [(11, 257), (7, 700), (9, 300)]
    data = {R: _sim(R, n, 1700 + R, kmax=3, min_sep=1.5, max_sep=4.0)[0] for R, n in sizes}
    ctx = Q.Context(_params(7))
    sets, outs = [], []
    for R, n in sizes:
        d_rois = torch.from_numpy(data[R].view(np.int16)).cuda()
        loc = torch.zeros((3 * n, 12), dtype=torch.float32, device="cuda")
        sel = torch.zeros(n, dtype=torch.int32, device="cuda")
        sets.append((R, d_rois, n, loc, sel))
        outs.append((R, n, d_rois, loc, sel))
    ctx.fit_mle_mixed_dev(sets, 3)
    torch.cuda.synchronize()
    for R, n, d_rois, loc, sel in outs:
        c1 = Q.Context(_params(R))
        l1 = torch.zeros((3 * n, 12), dtype=torch.float32, device="cuda")
        s1 = torch.zeros(n, dtype=torch.int32, device="cuda")
        c1.fit_mle_dev(d_rois, n, 3, l1, s1)
        torch.cuda.synchronize()
        print("R", R, "selected", np.bincount(s1.cpu().numpy(), minlength=4))
        np.testing.assert_array_equal(sel.cpu().numpy(), s1.cpu().numpy())
        np.testing.assert_array_equal(loc.cpu().numpy().view(np.uint32), l1.cpu().numpy().view(np.uint32))


def test_elliptical_model_refused_by_the_2d_only_paths(gpu):
    """The learned warm start and its labeller are trained on the 2D model and refuse AS3D."""
    R, n = 7, 64
    rois, _ = _sim(R, n, 1800)
    d_rois = torch.from_numpy(rois.view(np.int16)).cuda()
    ctx = Q.Context(_params(R))
    loc = torch.zeros((3 * n, 12), dtype=torch.float32, device="cuda")
    with pytest.raises(RuntimeError, match="2D model only"):
        ctx.fit_warm_dev(Q.Net(seed=1), d_rois, n, loc, None)
    with pytest.raises(RuntimeError, match="2D model only"):
        ctx.label_dev(d_rois, n, torch.zeros((n, 16), dtype=torch.float32, device="cuda"))


def test_elliptical_graph_replays_match_eager(gpu):
    """A captured plain-path batch of the elliptical model: every replay equals the eager call."""
    R, n = 9, 1201
    ctx = Q.Context(_params(R))
    buf = torch.from_numpy(_sim(R, n, 1900, kmax=3, min_sep=1.5, max_sep=4.0)[0].view(np.int16)).cuda()
    loc = torch.zeros((3 * n, 12), dtype=torch.float32, device="cuda")
    sel = torch.zeros(n, dtype=torch.int32, device="cuda")
    g = ctx.graph_mle(buf, n, 3, loc, sel)
    for rep, seed in enumerate((1901, 1902, 1901)):
        new = torch.from_numpy(_sim(R, n, seed, kmax=3, min_sep=1.5, max_sep=4.0)[0].view(np.int16)).cuda()
        buf.copy_(new)
        torch.cuda.synchronize()
        g.launch()
        torch.cuda.synchronize()
        got_l, got_s = loc.cpu().numpy(), sel.cpu().numpy()
        ref_l, ref_s = torch.zeros_like(loc), torch.zeros_like(sel)
        ctx.fit_mle_dev(new, n, 3, ref_l, ref_s)
        torch.cuda.synchronize()
        print("replay", rep, "counts", np.bincount(got_s, minlength=4))
        np.testing.assert_array_equal(got_s, ref_s.cpu().numpy())
        np.testing.assert_array_equal(got_l.view(np.uint32), ref_l.cpu().numpy().view(np.uint32))
    g.close()


def test_elliptical_autotune_tunes_the_kernel_choice_only(gpu):
    R, n = 9, 1 << 13
    ctx = Q.Context(_params(R))
    d_rois = torch.from_numpy(_sim(R, n, 2000, kmax=3, min_sep=1.5, max_sep=4.0)[0].view(np.int16)).cuda()
    tab = ctx.autotune_dev(d_rois, n, trials=1, min_select_agree=0.99, max_dx_p99=0.02)
    for e in tab:
        print(e)
    # one candidate per kernel choice: the persistent queue at all, 1, 2 and 4 waves per CU, and
    # the one-pass kernels; all give the reference's rows (the caps and stop are the model's own)
    assert len(tab) == 5 and not any(e["warm"] for e in tab)
    assert all(e["select_agree"] == 1.0 and e["dx_p99"] == 0.0 for e in tab)
    chosen = [e for e in tab if e["chosen"]]
    assert len(chosen) == 1 and chosen[0]["ms"] == min(e["ms"] for e in tab)
    c7 = Q.Context(_params(7))
    with pytest.raises(RuntimeError, match="2D model only"):
        c7.autotune_dev(torch.zeros((64, 56), dtype=torch.int16, device="cuda"), 64, net=Q.Net(seed=1))


def test_elliptical_pin_stop_option(gpu):
    """pin_stop (off by default) ends a three-emitter elliptical fit after 10 accepted iterations in
    a row at a box bound.  The ended fits are flagged and are points the selection rejects; every
    other fit is the default run's bit for bit; one- and two-emitter fits are untouched; the queue
    and one-pass kernels still agree bit for bit."""
    R, n = 9, 577
    rois = _sim(R, n, 2100, kmax=3, min_sep=1.0, max_sep=4.0, photons_min=800.0, photons_max=6000.0)[0]
    d_rois = torch.from_numpy(rois.view(np.int16)).cuda()
    ctx = Q.Context(_params(R))

    def chain():
        out, prev = [], None
        for K in (1, 2, 3):
            th, info, ll = _cfit(ctx, K, d_rois, n, Q.CFIT_INIT_HEUR if K == 1 else Q.CFIT_INIT_ADD, prev)
            prev = th
            out.append((th, info, ll))
        torch.cuda.synchronize()
        return [(t.cpu().numpy(), i.cpu().numpy(), l.cpu().numpy()) for t, i, l in out]

    assert ctx.tuning()["pin_stop"] == 0
    base = chain()
    ctx.set_tuning(pin_stop=1, queue=0)
    one = chain()
    ctx.set_tuning(queue=1)
    got = chain()
    for (t0, i0, l0), (t1, i1, l1) in zip(one, got):
        np.testing.assert_array_equal(t1.view(np.uint32), t0.view(np.uint32))
        np.testing.assert_array_equal(i1[:, :2], i0[:, :2])
        np.testing.assert_array_equal(l1.view(np.uint64), l0.view(np.uint64))
    for K in (1, 2):
        np.testing.assert_array_equal(got[K - 1][0].view(np.uint32), base[K - 1][0].view(np.uint32))
    t3, i3, _ = got[2]
    pinned = (i3[:, 1] & Q.CFIT_PINNED) != 0
    print("pinned", pinned.sum(), "of", n, "iterations", base[2][1][:, 0].mean(), "->", i3[:, 0].mean())
    assert pinned.sum() > 0 and i3[:, 0].mean() < base[2][1][:, 0].mean()
    assert not any(_valid(t3[i], 3, R) for i in np.nonzero(pinned)[0])
    np.testing.assert_array_equal(t3[~pinned].view(np.uint32), base[2][0][~pinned].view(np.uint32))
    with pytest.raises(RuntimeError, match="pin_stop"):
        ctx.set_tuning(pin_stop=2)
