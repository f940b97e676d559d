"""numpy restatement of the warm-start network (qc-storm_amd/csrc/nn.hip) for gradient checks.

Test infrastructure: float64 arithmetic with bf16 rounding at the points where the GPU stores
bf16 (inputs, weights, activations, back-propagated gradients), so the GPU's FP32-accumulated
MFMA results must agree to accumulation-order rounding."""
import numpy as np

NR, C1, C2, C3, F0, F1, NO, K1 = 7, 32, 64, 64, 80, 64, 16, 16
K2, K3 = 9 * C1, 9 * C2
CE_W = 4.0
SIZES = [("W1", C1 * K1), ("b1", C1), ("W2", C2 * K2), ("b2", C2), ("W3", C3 * K3), ("b3", C3), ("W4", F1 * F0),
         ("b4", F1), ("W5", NO * F1), ("b5", NO)]
SHAPES = {"W1": (C1, K1), "W2": (C2, K2), "W3": (C3, K3), "W4": (F1, F0), "W5": (NO, F1)}


def unpack(w):
    out, o = {}, 0
    for name, n in SIZES:
        v = w[o:o + n]
        out[name] = v.reshape(SHAPES[name]) if name in SHAPES else v
        o += n
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], np.float64).ravel() for name, _ in SIZES])


def bf(x):
    a = np.ascontiguousarray(x, np.float32).copy()
    nan = np.isnan(a)
    u = a.view(np.uint32)
    u += np.uint32(0x7FFF) + ((u >> np.uint32(16)) & np.uint32(1))
    u &= np.uint32(0xFFFF0000)
    a[nan] = np.nan
    return a.astype(np.float64)


def prep(rois, offset=100.0, kadc=0.45, qe=1.0):
    f32 = np.float32
    src = rois[:, :NR * NR].astype(f32)
    q = np.maximum((src - f32(offset)) * f32(kadc), f32(1.0)) * (f32(1.0) / f32(qe))
    q = q.astype(f32)
    mx = q.max(1)
    img = q.reshape(-1, NR, NR)
    e = [np.zeros(len(q), f32) for _ in range(4)]
    for k in range(NR):
        e[0] = (e[0] + img[:, 0, k]).astype(f32)
        e[1] = (e[1] + img[:, NR - 1, k]).astype(f32)
        e[2] = (e[2] + img[:, k, 0]).astype(f32)
        e[3] = (e[3] + img[:, k, NR - 1]).astype(f32)
    bh = (f32(0.5) * (np.minimum(e[0], e[1]) + np.minimum(e[2], e[3]))).astype(f32) / f32(NR)
    scale = (np.maximum(mx - bh, f32(0)) + f32(1)).astype(f32)
    inv = (f32(1) / scale).astype(f32)
    x = bf(((q - bh[:, None]) * inv[:, None]).astype(f32)).reshape(-1, NR, NR)
    side = np.zeros((len(q), F0 - C3))
    side[:, 0] = bf(np.log(scale.astype(np.float64)).astype(f32) * f32(0.1))
    side[:, 1] = bf((bh * inv).astype(f32))
    return x, side, scale.astype(np.float64), bh.astype(np.float64)


def forward(w, rois):
    P = unpack(w)
    Wb = {k: bf(P[k]) for k in SHAPES}
    x, side, scale, bh = prep(rois)
    n = len(x)
    x1 = np.zeros((n, 25, K1))
    for py in range(5):
        for px in range(5):
            for t in range(9):
                x1[:, py * 5 + px, t] = x[:, py + t // 3, px + t % 3]
    a1 = bf(np.maximum(x1 @ Wb["W1"].T + P["b1"], 0))                    # [n,25,32]
    x2 = np.zeros((n, 9, K2))
    for py in range(3):
        for px in range(3):
            for ty in range(3):
                for tx in range(3):
                    x2[:, py * 3 + px, (ty * 3 + tx) * C1:(ty * 3 + tx + 1) * C1] = a1[:, (py + ty) * 5 + px + tx, :]
    a2 = bf(np.maximum(x2 @ Wb["W2"].T + P["b2"], 0))                    # [n,9,64]
    a2f = a2.reshape(n, K3)
    a3 = np.zeros((n, F0))
    a3[:, :C3] = bf(np.maximum(a2f @ Wb["W3"].T + P["b3"], 0))
    a3[:, C3:] = side
    a4 = bf(np.maximum(a3 @ Wb["W4"].T + P["b4"], 0))
    out = a4 @ Wb["W5"].T + P["b5"]
    cache = dict(P=P, Wb=Wb, x1=x1, a1=a1, x2=x2, a2=a2, a2f=a2f, a3=a3, a4=a4, scale=scale, bh=bh)
    return out, cache


def loss_grad(out, labels, scale, bh):
    n = len(out)
    d = np.zeros_like(out)
    L = 0.0
    cnt = 0
    for i in range(n):
        K = int(labels[i, 0])
        if K < 1 or K > 3:
            continue
        cnt += 1
        o, lb = out[i], labels[i]
        z = o[:3] - o[:3].max()
        p = np.exp(z) / np.exp(z).sum()
        L += -CE_W * np.log(p[K - 1])
        d[i, :3] = CE_W * (p - np.eye(3)[K - 1])
        itgt = 1.0 / (2 * np.pi * scale[i])
        wp = NR * NR
        for k in range(K):
            ex = o[3 + 3 * k] - lb[1 + 3 * k] / NR
            ey = o[4 + 3 * k] - lb[2 + 3 * k] / NR
            en = o[5 + 3 * k] - lb[3 + 3 * k] * itgt
            L += wp * (ex * ex + ey * ey) + en * en
            d[i, 3 + 3 * k], d[i, 4 + 3 * k], d[i, 5 + 3 * k] = 2 * wp * ex, 2 * wp * ey, 2 * en
        es = o[12] - (lb[10] - 1.3)
        eb = o[13] - (lb[11] - bh[i]) / scale[i]
        L += es * es + 10 * eb * eb
        d[i, 12], d[i, 13] = 2 * es, 20 * eb
    return L / n, d / n


def backward(cache, d5):
    P, Wb = cache["P"], cache["Wb"]
    n = len(d5)
    g = {}
    g["b5"] = d5.sum(0)
    d5b = bf(d5)
    g["W5"] = d5b.T @ cache["a4"]
    da4 = d5b @ Wb["W5"]
    d4 = da4 * (cache["a4"] > 0)
    g["b4"] = d4.sum(0)
    d4b = bf(d4)
    g["W4"] = d4b.T @ cache["a3"]
    da3 = d4b @ Wb["W4"]
    d3 = da3[:, :C3] * (cache["a3"][:, :C3] > 0)
    g["b3"] = d3.sum(0)
    d3b = bf(d3)
    g["W3"] = d3b.T @ cache["a2f"]
    da2 = (d3b @ Wb["W3"]).reshape(n, 9, C2)
    d2 = da2 * (cache["a2"] > 0)
    g["b2"] = d2.sum((0, 1))
    d2b = bf(d2)
    g["W2"] = np.einsum("npo,npk->ok", d2b, cache["x2"])
    dx2 = d2b @ Wb["W2"]                                                 # [n,9,288]
    da1 = np.zeros((n, 25, C1))
    for py in range(3):
        for px in range(3):
            for ty in range(3):
                for tx in range(3):
                    da1[:, (py + ty) * 5 + px + tx, :] += dx2[:, py * 3 + px, (ty * 3 + tx) * C1:(ty * 3 + tx + 1) * C1]
    d1 = da1 * (cache["a1"] > 0)
    g["b1"] = d1.sum((0, 1))
    d1b = bf(d1)
    g["W1"] = np.einsum("npo,npk->ok", d1b, cache["x1"])
    g["W1"][:, 9:] = 0.0
    return pack(g)
