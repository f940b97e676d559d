"""The learned warm start on the GPU: MFMA GEMM layout, forward and gradient vs the numpy
restatement (tests/nn_ref.py), training on simulator ROIs labelled by the plain path, and the
warm-started fit against the plain path on held-out ROIs (RMSE against the simulated truth)."""
import ctypes as C

import numpy as np
import pytest
import torch

import nn_ref
import qcstorm as Q

pytestmark = pytest.mark.gpu


def _bf16_u16(a):
    return (nn_ref.bf(a).astype(np.float32).view(np.uint32) >> 16).astype(np.uint16)


def test_gemm_nt_bf16_matches_numpy(gpu):
    rng = np.random.default_rng(1)
    M, N, K = 80, 40, 48
    A = rng.normal(size=(M, K))
    B = rng.normal(size=(N, K))
    Au, Bu = _bf16_u16(A), _bf16_u16(B)
    dA = torch.from_numpy(Au.view(np.int16)).cuda()
    dB = torch.from_numpy(Bu.view(np.int16)).cuda()
    dC = torch.zeros((M, N), dtype=torch.float32, device="cuda")
    Q.check(Q.lib().qcs_dev_gemm_nt_bf16(C.c_void_p(dA.data_ptr()), C.c_void_p(dB.data_ptr()), C.c_void_p(dC.data_ptr()),
                                         M, N, K, None), "gemm")
    torch.cuda.synchronize()
    ref = nn_ref.bf(A) @ nn_ref.bf(B).T
    np.testing.assert_allclose(dC.cpu().numpy(), ref, rtol=1e-5, atol=1e-4)


def _sim(n, seed, kmax=3):
    sp = Q.default_sim(width=7, height=7, n_per_unit_max=kmax, min_sep=1.0, max_sep=4.0, photons_min=500.0,
                       photons_max=5000.0)
    d_rois = torch.empty((n, Q.roi_words(7)), dtype=torch.int16, device="cuda")
    d_truth = torch.empty((n, kmax, Q.TRUTH_FLOATS), dtype=torch.float32, device="cuda")
    Q.check(Q.lib().qcs_sim_rois_dev(C.byref(sp), seed, n, 1, C.c_void_p(d_rois.data_ptr()),
                                     C.c_void_p(d_truth.data_ptr()), None), "qcs_sim_rois_dev")
    torch.cuda.synchronize()
    return d_rois, d_truth


def _params():
    return Q.default_params(QE=1.0, ROISize=7, PixelSize=100.0)


def test_forward_and_gradient_match_reference(gpu):
    n = 256
    d_rois, _ = _sim(n, 11)
    p = _params()
    net = Q.Net(seed=3)
    w = net.params().astype(np.float64)
    out = torch.zeros((n, 16), dtype=torch.float32, device="cuda")
    net.forward_dev(p, d_rois, n, out)
    torch.cuda.synchronize()
    rois = d_rois.cpu().numpy().view(np.uint16)
    ref, cache = nn_ref.forward(w, rois)
    got = out.cpu().numpy()
    err = np.abs(got - ref).max()
    print("forward max |diff|", err, "output scale", np.abs(ref).max())
    assert err < 2e-2 * max(1.0, np.abs(ref).max())
    ctx = Q.Context(p)
    labels = torch.zeros((n, 16), dtype=torch.float32, device="cuda")
    ctx.label_dev(d_rois, n, labels)
    torch.cuda.synchronize()
    lab = labels.cpu().numpy()
    assert (lab[:, 0] > 0).mean() > 0.9
    g_gpu, loss_gpu = net.grad_dev(p, d_rois, labels, n)
    loss_ref, d5 = nn_ref.loss_grad(ref, lab, cache["scale"], cache["bh"])
    g_ref = nn_ref.backward(cache, d5)
    print("loss gpu/ref", loss_gpu, loss_ref)
    assert abs(loss_gpu - loss_ref) < 1e-3 * max(1.0, abs(loss_ref))
    o = 0
    for name, size in nn_ref.SIZES:
        a, b = g_gpu[o:o + size].astype(np.float64), g_ref[o:o + size]
        o += size
        nb = np.linalg.norm(b)
        rel = np.linalg.norm(a - b) / max(nb, 1e-12)
        cos = float(a @ b / max(np.linalg.norm(a) * nb, 1e-30))
        print(name, "rel", rel, "cos", cos)
        assert rel < 3e-2 and cos > 0.999


@pytest.mark.parametrize("n", [1, 1003, 70001])
def test_fused_forward_matches_layer_by_layer(gpu, n):
    """The fused inference kernel (transposed MFMA chain, activations in registers) against the
    training path's layer-by-layer GEMMs: same k-chunk order and rounding points."""
    d_rois, _ = _sim(n, 21 + n)
    p = _params()
    net = Q.Net(seed=9)
    o_f = torch.zeros((n, 16), dtype=torch.float32, device="cuda")
    o_l = torch.zeros((n, 16), dtype=torch.float32, device="cuda")
    m_f = torch.zeros((n, 2), dtype=torch.float32, device="cuda")
    m_l = torch.zeros((n, 2), dtype=torch.float32, device="cuda")
    net.forward_dev(p, d_rois, n, o_f, m_f)
    net.forward_dev(p, d_rois, n, o_l, m_l, layers=True)
    torch.cuda.synchronize()
    a, b = o_f.cpu().numpy(), o_l.cpu().numpy()
    print("n", n, "max |fused - layers|", np.abs(a - b).max(), "bitwise rows", (a.view(np.uint32) == b.view(np.uint32)).all(1).mean())
    np.testing.assert_array_equal(m_f.cpu().numpy().view(np.uint32), m_l.cpu().numpy().view(np.uint32))
    np.testing.assert_array_equal(a.view(np.uint32), b.view(np.uint32))


@pytest.fixture(scope="module")
def trained():
    """A network trained on 2^23 labelled simulator ROIs (2048 Adam steps of 4096, cosine decay)."""
    from qcstorm import warmstart as WS
    net, hist = WS.train(n_rois=1 << 23, batch=4096, epochs=1, lr=2e-3, seed=2004, log=print)
    losses = [l for _, l in hist["loss"]]
    print("training losses", [round(x, 3) for x in losses], "seconds", hist["seconds"])
    return net, losses


def test_training_reduces_loss_and_predicts_counts(gpu, trained):
    net, losses = trained
    assert losses[-1] < 0.35 * losses[0]
    p = _params()
    n = 8192
    d_rois, d_truth = _sim(n, 6001)
    ctx = Q.Context(p)
    labels = torch.zeros((n, 16), dtype=torch.float32, device="cuda")
    ctx.label_dev(d_rois, n, labels)
    out = torch.zeros((n, 16), dtype=torch.float32, device="cuda")
    net.forward_dev(p, d_rois, n, out)
    torch.cuda.synchronize()
    k_pred = out[:, :3].argmax(1).cpu().numpy() + 1
    k_lab = labels[:, 0].cpu().numpy().astype(int)
    ok = k_lab > 0
    acc = (k_pred[ok] == k_lab[ok]).mean()
    print("count accuracy vs plain-path labels", acc)
    assert acc > 0.78


def _match_rmse(loc, truth, n):
    """Greedy nearest matching (within 1 px) of fitted rows to simulated emitters; returns
    (rmse px, recall, precision)."""
    loc = loc.reshape(n, 3, 12)
    d2, tp, nt, nf = [], 0, 0, 0
    for i in range(n):
        t = truth[i][truth[i, :, 4] > 0]
        f = loc[i][loc[i, :, 0] > 0]
        nt += len(t)
        nf += len(f)
        used = set()
        for e in t:
            best, bj = 1.0, -1
            for j, r in enumerate(f):
                if j in used:
                    continue
                dd = (r[1] - 100 + 3 - e[1]) ** 2 + (r[2] - 100 + 3 - e[2]) ** 2
                if dd < best:
                    best, bj = dd, j
            if bj >= 0:
                used.add(bj)
                d2.append(best)
                tp += 1
    return float(np.sqrt(np.mean(d2))), tp / max(nt, 1), tp / max(nf, 1)


def test_warm_path_matches_plain_accuracy(gpu, trained):
    net, _ = trained
    p = _params()
    n = 16384
    d_rois, d_truth = _sim(n, 7001)
    truth = d_truth.cpu().numpy()
    ctx = Q.Context(p)
    loc_p = torch.zeros((3 * n, 12), dtype=torch.float32, device="cuda")
    loc_w = torch.zeros((3 * n, 12), dtype=torch.float32, device="cuda")
    stp = ctx.fit_plain_stats_dev(d_rois, n, 3, loc_p)
    st = ctx.fit_warm_dev(net, d_rois, n, loc_w)
    torch.cuda.synchronize()
    # timing of both paths (events inside the library)
    ctx.timing_enable(True)
    ctx.fit_mle_dev(d_rois, n, 3, loc_p)
    tp = ctx.timing()
    ctx.timing_enable(True)
    ctx.fit_warm_dev(net, d_rois, n, loc_w, stats=False)
    tw = ctx.timing()
    plain_ms, warm_ms = tp["cfit_ms"], tw["cfit_ms"] + tw["nn_ms"]
    rp = _match_rmse(loc_p.cpu().numpy(), truth, n)
    rw = _match_rmse(loc_w.cpu().numpy(), truth, n)
    print("stats", st, "plain ms %.2f warm ms %.2f (nn %.2f) speedup %.2f" % (plain_ms, warm_ms, tw["nn_ms"],
                                                                           plain_ms / warm_ms))
    print("plain rmse/recall/precision", rp, "warm", rw)
    print("BFGS iterations per ROI: plain %.2f warm %.2f" % (stp["iters"] / n, st["iters"] / n))
    assert st["fallbacks"] < 0.5 * n
    assert rw[0] <= 1.03 * rp[0]
    assert rw[1] >= rp[1] - 0.02
    # the measured gain of the network's starts and count caps: at least a fifth fewer BFGS
    # iterations per ROI than the plain path over the same ROIs (about 40% fewer on this data)
    assert st["iters"] < 0.8 * stp["iters"]
