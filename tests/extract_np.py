"""Independent numpy formulation of the extraction filters (test infrastructure).

Vectorised restatement of SE/LDROIExtraction_ImageFiltering.cu (background removal, separable
smoothing) and SE/LDROIExtraction_ROIExtraction.cu:25-164 (local-maximum test), used to pin the C
oracle.  Float32 arithmetic; the multiply-add of the smoothing is emulated exactly through
float64 (products of float32 operands are exact in float64).
"""
import numpy as np


def _clamp_take(img, ys, xs):
    h, w = img.shape
    return img[np.clip(ys, 0, h - 1)[:, None], np.clip(xs, 0, w - 1)[None, :]]


def remove_background(img):
    img = img.astype(np.int64)
    bh, w = img.shape
    wt = (w + 31) // 32 * 32
    ys = np.arange(bh)
    xs = np.arange(wt)
    s = np.zeros((bh, wt), np.int64)
    for i in range(8):
        l1 = _clamp_take(img, ys - 4, xs - 4 + i)
        l2 = _clamp_take(img, ys - 4 + i, xs + 4)
        l3 = _clamp_take(img, ys + 4, xs - 3 + (7 - i))
        l4 = _clamp_take(img, ys - 3 + (7 - i), xs - 4)
        s += np.minimum(np.minimum(l1, l2), np.minimum(l3, l4))
    b = (s.astype(np.float32) / np.float32(8.0)).astype(np.float32)
    x = np.arange(w)
    g0 = x & ~31
    xl = np.maximum(x - 1, g0)
    xr = np.minimum(x + 1, g0 + 31)
    mb = ((b[:, xl] + b[:, x]) + b[:, xr]) / np.float32(3.0)
    v = img[:, :w].astype(np.float32) - mb.astype(np.float32)
    v = np.maximum(v, np.float32(0.0))
    return v.astype(np.uint16)


def _fma32(a, b, c):
    return (a.astype(np.float64) * b.astype(np.float64) + c.astype(np.float64)).astype(np.float32)


def conv_h(img, taps):
    bh, w = img.shape
    acc = np.zeros((bh, w), np.float32)
    for k in range(5):
        xs = np.clip(np.arange(w) - 2 + k, 0, w - 1)
        acc = _fma32(img[:, xs].astype(np.float32), np.float32(taps[k]), acc)
    return np.minimum(acc, 65535).astype(np.uint16)


def conv_v(img, taps):
    return np.ascontiguousarray(conv_h(np.ascontiguousarray(img.T), taps).T)


def find(det, h, variance, R, th_hi, th_lo):
    sd = np.sqrt(np.float32(variance)).astype(np.float32)
    th, th1 = np.float32(th_hi) * sd, np.float32(th_lo) * sd
    bh, w = det.shape
    d = det.astype(np.int64)
    half = R // 2
    pad = np.pad(d, 1)
    c = d
    nb = {(dy, dx): pad[1 + dy:1 + dy + bh, 1 + dx:1 + dx + w] for dy in (-1, 0, 1) for dx in (-1, 0, 1)}
    s8 = sum((nb[k].astype(np.float32) > th1).astype(int) for k in nb if k != (0, 0))
    si = (c > nb[(-1, 0)]).astype(int) + (c >= nb[(0, 1)]) + (c >= nb[(1, 0)]) + (c > nb[(0, -1)])
    so = (c > nb[(-1, -1)]).astype(int) + (c > nb[(-1, 1)]) + (c > nb[(1, -1)]) + (c > nb[(1, 1)])
    y = np.arange(bh)[:, None] % h
    x = np.arange(w)[None, :]
    inside = (x >= half) & (y >= half) & (x < w - half) & (y < h - half)
    ok = inside & (c.astype(np.float32) >= th) & (s8 >= 7) & (si == 4) & (so >= 3)
    cand = ok & (so == 4)
    ys, xs = np.nonzero(cand)
    return ok, np.stack([xs, ys], 1)
