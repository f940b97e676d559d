"""Synthetic localization rows of a fixed structure under a known drift (drift-correction tests).

Structure points are localized in every frame where they are on (probability p_on), at the point
plus the frame's drift plus localization noise; rows are sorted by frame."""
import numpy as np


def drifting_rows(W, H, nframes, npts, seed, ax=0.01, ay=-0.006, wobble=0.15, p_on=0.5, noise=0.03, margin=6.0):
    rng = np.random.default_rng(seed)
    pts = np.stack([rng.uniform(margin, W - margin, npts), rng.uniform(margin, H - margin, npts)], 1)
    f = np.arange(1, nframes + 1, dtype=np.float64)
    dx = ax * f + wobble * np.sin(f / 9.0)
    dy = ay * f + wobble * np.cos(f / 7.0) - wobble
    rows = []
    for k in range(nframes):
        on = rng.random(npts) < p_on
        m = int(on.sum())
        r = np.zeros((m, 12), np.float32)
        r[:, 0] = rng.uniform(200, 2000, m)
        r[:, 1] = pts[on, 0] + dx[k] + rng.normal(0, noise, m)
        r[:, 2] = pts[on, 1] + dy[k] + rng.normal(0, noise, m)
        r[:, 4:6] = 1.3
        r[:, 6] = r[:, 0] * 6
        r[:, 7] = 20
        r[:, 8] = rng.uniform(8, 40, m)
        r[:, 9] = rng.uniform(6, 14, m)
        r[:, 10] = r[:, 9]
        r[:, 11] = k + 1
        rows.append(r)
    return np.concatenate(rows), dx.astype(np.float32), dy.astype(np.float32), pts
