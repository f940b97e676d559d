"""Converged multi-emitter fitting on the GPU vs the FP64 CPU oracle.

The GPU fitter (BFGS, FP32, analytic gradient) and the oracle's FP64 Levenberg-Marquardt MLE
(oracle/orc_truth.c) maximise the same Poisson likelihood; started from the GPU result the
oracle must stay within the FP32 convergence tolerance.  The FP64 full-Fisher CRLB written by
qcs_fit_mle_dev must equal the oracle's at the same parameters, and the FP64 likelihood-ratio
model selection must recover the simulated emitter counts."""
import numpy as np
import pytest
import torch

import oracle_lib as O
import qcstorm as Q

pytestmark = pytest.mark.gpu


def _q(rois, R, p):
    img = rois[:, :R * R].astype(np.float64).reshape(-1, R, R)
    return np.maximum((img - p.Offset) * np.float32(p.KAdc), 1.0) / p.QE


def _cfit(ctx, K, rois, n, mode=Q.CFIT_INIT_HEUR, init=None):
    P = 3 * K + 2
    d_rois = torch.from_numpy(rois.view(np.int16)).cuda()
    th = torch.zeros((n, P), dtype=torch.float32, device="cuda")
    info = torch.zeros((n, 4), dtype=torch.int32, device="cuda")
    ll = torch.zeros(n, dtype=torch.float64, device="cuda")
    d_init = torch.from_numpy(np.ascontiguousarray(init, np.float32)).cuda() if init is not None else None
    ctx.cfit_dev(K, d_rois, n, mode, d_init, th, info, ll)
    torch.cuda.synchronize()
    return th.cpu().numpy(), info.cpu().numpy(), ll.cpu().numpy()


@pytest.mark.parametrize("R", [7, 9, 11])
def test_single_emitter_matches_fp64_oracle(gpu, R):
    n = 512
    sp = Q.default_sim(width=R, height=R)
    rois, truth = Q.sim_rois_host(sp, 900 + R, n)
    p = Q.default_params(QE=1.0, ROISize=R)
    ctx = Q.Context(p)
    th, info, ll = _cfit(ctx, 1, rois, n)
    conv = (info[:, 1] & Q.CFIT_CONVERGED) != 0
    print("R", R, "converged", conv.mean(), "iters p50/p95", np.percentile(info[:, 0], [50, 95]))
    assert conv.mean() > 0.97
    q = _q(rois, R, p)
    dx, dn, dll = [], [], []
    for i in np.nonzero(conv)[0][:200]:
        t0 = np.array([th[i, 0], th[i, 1], th[i, 2], th[i, 3], th[i, 4]], np.float64)
        t, it, l64 = O.mle(q[i], K=1, theta0=t0)
        dx.append(max(abs(t[1] - th[i, 1]), abs(t[2] - th[i, 2])))
        dn.append(abs(t[0] - th[i, 0]) / t[0])
        dll.append(abs(l64 - ll[i]))
    dx, dn, dll = np.array(dx), np.array(dn), np.array(dll)
    print("max |dx| p50/p99/max", np.percentile(dx, [50, 99]), dx.max(), "photons rel p99", np.percentile(dn, 99),
          "loglik diff max", dll.max())
    assert np.percentile(dx, 99) < 1e-3 and dx.max() < 5e-3
    assert np.percentile(dn, 99) < 1e-3
    assert dll.max() < 1e-2   # FP64 log-likelihood at nearly the same optimum
    # the fit tracks the simulated truth
    err = np.hypot(th[conv, 1] - truth[conv, 0, 1], th[conv, 2] - truth[conv, 0, 2])
    assert np.median(err) < 0.1


def test_two_emitters_from_residual_init_match_fp64_oracle(gpu):
    R, n = 9, 512
    sp = Q.default_sim(width=R, height=R, n_per_unit_max=2, min_sep=2.5, max_sep=4.0, photons_min=1500.0,
                       photons_max=5000.0)
    rois, truth = Q.sim_rois_host(sp, 31, n)
    p = Q.default_params(QE=1.0, ROISize=R)
    ctx = Q.Context(p)
    t1, _, _ = _cfit(ctx, 1, rois, n)
    t2, info, ll = _cfit(ctx, 2, rois, n, Q.CFIT_INIT_ADD, t1)
    two = truth[:, 1, 4] > 0
    conv = ((info[:, 1] & Q.CFIT_CONVERGED) != 0) & two
    print("two-emitter ROIs", two.sum(), "converged", conv.sum())
    assert conv.sum() > 0.9 * two.sum()
    q = _q(rois, R, p)
    dx = []
    for i in np.nonzero(conv)[0][:150]:
        t, it, l64 = O.mle(q[i], K=2, theta0=t2[i].astype(np.float64))
        dx.append(np.abs(t[[1, 2, 4, 5]] - t2[i, [1, 2, 4, 5]]).max())
    dx = np.array(dx)
    print("2E max |dx| p50/p95", np.percentile(dx, [50, 95]))
    assert np.percentile(dx, 95) < 2e-3


def test_fit_mle_selection_and_fp64_crlb(gpu):
    R, n = 9, 768
    sp = Q.default_sim(width=R, height=R, n_per_unit_max=3, min_sep=2.5, max_sep=4.0, photons_min=2000.0,
                       photons_max=6000.0, bg_min=10.0, bg_max=40.0)
    rois, truth = Q.sim_rois_host(sp, 77, n)
    p = Q.default_params(QE=1.0, ROISize=R, PixelSize=100.0)
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
    assert acc[1] > 0.9 and acc[2] > 0.8 and acc[3] > 0.5
    # rows: TotPho is N, PPho = N / (2 pi s^2); CRLB equals the oracle's full-Fisher bound
    checked = 0
    for i in range(n):
        K = sel[i]
        if K == 0:
            continue
        rows = loc[i, :K]
        s = rows[0, 4]
        th = []
        for k in range(K):
            th += [rows[k, 6], rows[k, 1] - 100 + R // 2, rows[k, 2] - 100 + R // 2]
        th += [s, rows[0, 7]]
        cr = O.fisher_crlb(R, np.array(th, np.float64), K=K)
        for k in range(K):
            np.testing.assert_allclose(rows[k, 9], cr[3 * k + 1] * 100.0, rtol=2e-5)
            np.testing.assert_allclose(rows[k, 10], cr[3 * k + 2] * 100.0, rtol=2e-5)
            np.testing.assert_allclose(rows[k, 0], rows[k, 6] / (2 * np.pi * s * s), rtol=1e-5)
        checked += 1
        if checked >= 150:
            break
    assert checked >= 150
