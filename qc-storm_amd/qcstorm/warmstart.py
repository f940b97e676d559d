"""Train the learned warm start on the GPU and save its parameters.

The training set is simulator ROIs (7x7, 1-3 emitters 1-4 px apart, photons log-U(500, 5000),
background U(10, 100), Poisson noise, ADU = e/0.45 + 100) labelled by the plain converged path
(qcs_label_dev: K = 1..3 fits and FP64 likelihood-ratio model selection).  Training runs the
library's own forward/backward kernels (BF16 MFMA) with Adam.

    python -m qcstorm.warmstart --rois 10000000 --out qc-storm_amd/weights/warmstart_r7.npy
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

from . import Context, Net, check, default_params, default_sim, lib, roi_words, TRUTH_FLOATS

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_WEIGHTS = os.path.join(os.path.dirname(HERE), "weights", "warmstart_r7.npy")


def sim_dev(n, seed, kmax=3, first_frame=1):
    """Simulator ROIs on the GPU: (rois int16 [n][56], truth float32 [n][kmax][8])."""
    import torch
    sp = default_sim(width=7, height=7, n_per_unit_max=kmax, min_sep=1.0, max_sep=4.0, photons_min=500.0,
                     photons_max=5000.0)
    rois = torch.empty((n, roi_words(7)), dtype=torch.int16, device="cuda")
    truth = torch.empty((n, kmax, TRUTH_FLOATS), dtype=torch.float32, device="cuda")
    check(lib().qcs_sim_rois_dev(C.byref(sp), seed, n, first_frame, C.c_void_p(rois.data_ptr()),
                                 C.c_void_p(truth.data_ptr()), None), "qcs_sim_rois_dev")
    return rois, truth


def params():
    return default_params(QE=1.0, ROISize=7, PixelSize=100.0)


def train(n_rois=1 << 20, batch=2048, epochs=1, lr=2e-3, seed=2004, net_seed=7, chunk=1 << 20, log=print):
    """Train a fresh network on n_rois labelled ROIs; returns (net, history)."""
    import torch
    p = params()
    ctx = Context(p)
    net = Net(seed=net_seed)
    hist = {"loss": [], "steps": 0, "n_rois": n_rois, "batch": batch, "epochs": epochs, "lr": lr, "seed": seed}
    t0 = time.time()
    step = 0
    total_steps = epochs * ((n_rois + batch - 1) // batch)
    for ep in range(epochs):
        for c0 in range(0, n_rois, chunk):
            m = min(chunk, n_rois - c0)
            rois, _ = sim_dev(m, seed + 1000 * ep + c0 // chunk)   # fresh ROIs per chunk and epoch
            labels = torch.zeros((m, 16), dtype=torch.float32, device="cuda")
            ctx.label_dev(rois, m, labels)
            for b0 in range(0, m - batch + 1, batch):
                # cosine decay to lr / 20 over the run
                f = 0.05 + 0.95 * 0.5 * (1.0 + np.cos(np.pi * step / max(total_steps, 1)))
                want = step % 200 == 0
                loss = net.train_step_dev(p, rois[b0:b0 + batch], labels[b0:b0 + batch], batch, lr * f, want_loss=want)
                if want:
                    hist["loss"].append((step, loss))
                    log("step %d loss %.4f (%.1f s)" % (step, loss, time.time() - t0))
                step += 1
    torch.cuda.synchronize()
    hist["steps"] = step
    hist["seconds"] = time.time() - t0
    ctx.close()
    return net, hist


def load(path=DEFAULT_WEIGHTS, seed=0):
    """A network with the saved parameters (raises if the file is missing)."""
    net = Net(seed=seed)
    net.set_params(np.load(path, allow_pickle=False))
    return net


def main():
    ap = argparse.ArgumentParser(description="train the warm-start network")
    ap.add_argument("--rois", type=int, default=10_000_000)
    ap.add_argument("--batch", type=int, default=2048)
    ap.add_argument("--epochs", type=int, default=1)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--seed", type=int, default=2004)
    ap.add_argument("--out", default=DEFAULT_WEIGHTS)
    args = ap.parse_args()
    net, hist = train(args.rois, args.batch, args.epochs, args.lr, args.seed,
                      log=lambda s: print(s, file=sys.stderr, flush=True))
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    np.save(args.out, net.params())
    with open(os.path.splitext(args.out)[0] + ".json", "w") as f:
        json.dump(hist, f)
    print(json.dumps({"saved": args.out, "steps": hist["steps"], "seconds": hist["seconds"],
                      "final_loss": hist["loss"][-1] if hist["loss"] else None}))


if __name__ == "__main__":
    main()
