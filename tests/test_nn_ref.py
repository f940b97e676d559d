"""The numpy restatement of the warm-start network (the reference of the GPU gradient check) is
itself pinned: with the bf16 roundings switched off, its backward pass equals central finite
differences of its loss."""
import numpy as np

import nn_ref
import qcstorm as Q


def test_reference_backward_matches_finite_differences(monkeypatch):
    monkeypatch.setattr(nn_ref, "bf", lambda x: np.asarray(x, np.float64))
    sp = Q.default_sim(width=7, height=7, n_per_unit_max=3)
    rois, _ = Q.sim_rois_host(sp, 5, 6)
    rng = np.random.default_rng(0)
    w = rng.normal(size=sum(s for _, s in nn_ref.SIZES)) * 0.05
    P = nn_ref.unpack(w)
    P["W1"][:, 9:] = 0
    w = nn_ref.pack(P)
    lab = np.zeros((6, 16))
    lab[:, 0] = [1, 2, 3, 1, 2, 3]
    for i in range(6):
        for k in range(3):
            lab[i, 1 + 3 * k:4 + 3 * k] = [3.2 + 0.3 * k, 3.6 - 0.2 * k, 1500 - 300 * k]
        lab[i, 10], lab[i, 11] = 1.4, 40.0

    def loss(wv):
        out, c = nn_ref.forward(wv, rois)
        return nn_ref.loss_grad(out, lab, c["scale"], c["bh"])[0]

    out, c = nn_ref.forward(w, rois)
    _, d5 = nn_ref.loss_grad(out, lab, c["scale"], c["bh"])
    g = nn_ref.backward(c, d5)
    o, w1 = 0, None
    for name, s in nn_ref.SIZES:
        if name == "W1":
            w1 = (o, o + s)
        o += s
    errs = []
    for i in rng.choice(len(w), 60, replace=False):
        if w1[0] <= i < w1[1] and (i - w1[0]) % 16 >= 9:
            continue
        wp, wm = w.copy(), w.copy()
        wp[i] += 1e-6
        wm[i] -= 1e-6
        fd = (loss(wp) - loss(wm)) / 2e-6
        errs.append(abs(fd - g[i]) / max(1e-8, abs(fd) + abs(g[i])))
    assert max(errs) < 1e-4


def test_bf16_rounding_is_nearest_even():
    x = np.array([1.0, 1.00390625, 1.005859375, -2.5, 3.0e-39, np.inf], np.float32)
    r = nn_ref.bf(x)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0078125 and r[3] == -2.5 and np.isinf(r[5])
