#!/usr/bin/env python3
"""Throughput of the MI355X localization path on the full-frame 2D workload.

Workload: 200 uint16 frames of 2048 x 2048 px (100 nm px), emitters uniform at 0.3 /um^2
(0.003 per px, ~12.6k per frame), PSF sigma U(1.0, 1.6) px, photons log-U(500, 5000), background
U(10, 100) photons/px, Poisson noise, ADU = e/0.45 + 100, QE = 1 (simulator seed 1001).  The frames
are simulated on the GPU and stay resident in HBM.  One step is one pass of the hot path over the
200 frames, all on one HIP stream:

    ROI extraction (qcs_extract_dev, batches of 1 frame, 7x7 ROIs + WLE pre-estimation)
    -> single-emitter 2D BFGS MLE fit + CRLB (qcs_fit_dev)
    -> frame-sorted compaction (qcs_compact_sorted_dev)
    -> 10 nm-pixel rendering into a 20480 x 20480 float image, a fresh image per step
       (qcs_render_fresh_dev: the same image as clearing it and calling qcs_render_dev, drawn in
       one write pass)

`value` = fitted ROIs per second over whole steps (every extracted ROI is fitted).  The roofline
entry is for the dominant kernel (the 7x7 one-emitter BFGS fit): algorithmic FLOPs per launch
(120,736 FLOP per fit = 176 loss evaluations x 49 pixels x 14 FLOP, FMA = 2) divided by the
average launch duration, measured with HIP events recorded on the launch stream by the library.
The CPU baseline is the oracle's OpenMP fit + CRLB (same schedule, bit-identical output) on a
bounded sample of the same ROIs, on the host cores.  `parity_sample` counts the bit-exact rows of
that sample and (`vs_fp64_mle`) states the rows' error against the oracle's FP64 maximum-
likelihood fit run to convergence, on the first 16,384 kept rows, separately for unweighted
(MLEFit) and WLE-weighted ROIs.

A second leg (`multi_emitter` in the JSON line) fits 2^20 simulator 7x7 ROIs with 1-3 emitters
1-4 px apart twice: the plain converged path (K = 1, 2, 3 fits and FP64 likelihood-ratio model
selection) and the learned warm start (network -> converged fit of the predicted count, plain
fallback), reporting both throughputs, the speedup, and RMSE / recall / RMSE-over-CRLB against
the simulated truth on the first 65,536 ROIs.  A third (`as3d_multi_emitter`) fits 2^18 astigmatic
9x9 ROIs with 1-3 emitters with the elliptical model (per-emitter widths, up to 16 parameters) and
reports fits/s, the count accuracy, RMSE over CRLB and the z error against the simulated truth.

`hd_frames`: 1000 high-density 1024 x 1024 frames (2-5 emitters / um^2 with planted pairs and
triples) through the frames -> rows path of the converged fitter (mixed 7/9/11 ROIs, persistent
work-queue fits, FP64 selection and CRLB), one HIP graph replay per 50-frame batch; frames/s,
ROIs/s, accuracy of the first batch against the simulated emitters, and the learned warm start
against the plain path on the 7 x 7 ROIs those frames yield (median of 5 timed calls each).
`as3d_frames`: 500 astigmatic-3D 1024 x 1024 frames (9 x 9 ROIs, batches of 4 frames) through the
same pipeline as the headline step, frames/s.
`rois_100k`: the 100k seeded 7 x 7 ROIs of the single-emitter reference case, the oracle on every
usable host core against the library, rows compared bit for bit.  `session_path`: the plug-ins'
drop-in path, host frames through the online session with the result file, H2D included.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--math exact|fast] [--no-cpu-baseline]
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "qc-storm_amd"))

import numpy as np  # noqa: E402

import qcstorm as Q  # noqa: E402

FLOP_PER_FIT_7X7 = 176 * 49 * 14          # loss evaluations x pixels x FLOP per pixel-evaluation
PEAK_FP32_TFLOPS = 157.3                  # MI355X FP32 vector peak (dense)
METRIC = "ROI fits/s (7×7 2D & multi-emitter) + warm-start speedup at equal RMSE"
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "fit_gs2d_r7_fullframe_traffic.txt")


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def traffic_per_fit():
    """HBM bytes per fitted ROI of the fit kernel, from the committed PMC summary (or None)."""
    try:
        with open(TRAFFIC_FILE) as f:
            for line in f:
                if line.startswith("bytes_per_fit:"):
                    return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def host_cpu():
    """(usable host cores, nproc, CPU model) of the machine running the benchmark."""
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return usable, os.cpu_count() or 1, model


def cpu_baseline(params, rois, wle, seconds):
    """Oracle fit + CRLB (OpenMP over ROIs, every usable host core) on a bounded sample; returns
    (dict, sample rows)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O  # the oracle is used here only as the CPU baseline and checker
    threads, nproc, model = host_cpu()
    pilot = min(rois.shape[0], 4096 * threads)
    t0 = time.perf_counter()
    O.fit_omp(params, rois[:pilot], wle[:pilot], threads)
    rate = pilot / max(time.perf_counter() - t0, 1e-6)
    n = int(min(rois.shape[0], max(pilot, rate * seconds)))
    t0 = time.perf_counter()
    loc = O.fit_omp(params, rois[:n], wle[:n], threads)
    dt = time.perf_counter() - t0
    return ({"value": n / dt, "unit": "fits/s", "cores": threads, "kind": "port", "nproc": nproc, "cpu_model": model,
             "per_core_fits_per_s": n / dt / threads,
             "sample": "fit+CRLB of the first %d extracted 7x7 ROIs of the workload (%.1f s, %d OpenMP threads)"
                       % (n, dt, threads)}, loc)


def c0_leg(log, n=100000, seed=1000):
    """The 100k-ROI single-emitter case on the reference CPU path: n seeded 7 x 7 ROIs (one
    Gaussian each, centre within +-1 px, sigma U(1.0, 1.6), photons log-U(500, 5000), background
    U(10, 100), Poisson noise, ADU = e/0.45 + 100), WLE then fit + CRLB: the oracle on every usable
    host core against the library (host-pointer call, PCIe copies included, and the device call
    on resident ROIs), with the rows compared bit for bit."""
    import torch
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O   # CPU baseline and checker only
    sp = Q.default_sim()
    rois, _ = Q.sim_rois_host(sp, seed, n)
    p = Q.default_params(QE=1.0)
    wle = O.wle(rois, 7)
    threads, nproc, model = host_cpu()
    O.fit_omp(p, rois[:4096], wle[:4096], threads)
    t0 = time.perf_counter()
    ref = O.fit_omp(p, rois, wle, threads)
    cpu_s = time.perf_counter() - t0
    ctx = Q.Context(p)
    got = ctx.fit_rois(rois, wle)     # warm-up, also the parity rows
    t0 = time.perf_counter()
    got = ctx.fit_rois(rois, wle)
    host_s = time.perf_counter() - t0
    d_rois = torch.from_numpy(rois.view(np.int16)).cuda()
    d_wle = torch.from_numpy(wle).cuda()
    d_loc = torch.empty((n, 12), dtype=torch.float32, device="cuda")
    ctx.fit_dev(d_rois, d_wle, n, d_loc)
    torch.cuda.synchronize()
    ts = []
    for _ in range(5):
        t0 = time.perf_counter()
        ctx.fit_dev(d_rois, d_wle, n, d_loc)
        torch.cuda.synchronize()
        ts.append(time.perf_counter() - t0)
    ctx.close()
    same = int(np.all(got.view(np.uint32) == ref.view(np.uint32), axis=1).sum())
    log("bench: 100k-ROI case: CPU %.3g fits/s (%d threads), GPU %.3g fits/s" % (n / cpu_s, threads, n / np.median(ts)))
    return {"rois": n, "seed": seed, "cpu": {"value": n / cpu_s, "unit": "fits/s", "cores": threads, "nproc": nproc,
                                            "cpu_model": model, "kind": "port", "seconds": cpu_s},
            "gpu_host_pointer_fits_per_s": n / host_s, "gpu_device_fits_per_s_median": n / float(np.median(ts)),
            "rows_bit_exact": same, "rows": n, "accepted": int((ref[:, 0] > 0).sum())}


def drift_workload(W, H, nframes, npts, seed, ax=0.005, ay=-0.003, wobble=0.1):
    """Frame-sorted rows of a fixed structure (npts points, each on in half of the frames)
    localized under a known drift: (rows, drift x per frame, drift y per frame)."""
    rng = np.random.default_rng(seed)
    pts = np.stack([rng.uniform(8, W - 8, npts), rng.uniform(8, H - 8, npts)], 1).astype(np.float32)
    f = np.arange(1, nframes + 1, dtype=np.float64)
    dx = (ax * f + wobble * np.sin(f / 25.0)).astype(np.float32)
    dy = (ay * f + wobble * (np.cos(f / 19.0) - 1.0)).astype(np.float32)
    out = []
    for k in range(nframes):
        on = rng.random(npts) < 0.5
        m = int(on.sum())
        r = np.zeros((m, 12), np.float32)
        r[:, 0] = rng.uniform(200, 2000, m)
        r[:, 1] = pts[on, 0] + dx[k] + rng.normal(0, 0.03, m).astype(np.float32)
        r[:, 2] = pts[on, 1] + dy[k] + rng.normal(0, 0.03, m).astype(np.float32)
        r[:, 4:6] = 1.3
        r[:, 6] = r[:, 0] * 6
        r[:, 7] = 20
        r[:, 8] = rng.uniform(8, 40, m)
        r[:, 9] = rng.uniform(6, 14, m)
        r[:, 10] = r[:, 9]
        r[:, 11] = k + 1
        out.append(r)
    return np.concatenate(out), dx, dy


def match_rmse(loc, truth, pixel_size, R=7, zerr=False):
    """Greedy nearest matching (within 1 px) of fitted rows [n][3][12] to simulated emitters
    [n][3][8] of R x R ROIs (header Xc = Yc = 100); returns rmse (px, 2D: sqrt of mean dx^2 + dy^2),
    recall, precision and that RMSE over the bound sqrt(mean CRLBx^2 + CRLBy^2) of the matched rows
    (and, with zerr, the z RMSE and median |dz| of the matched emitters, nm)."""
    n = loc.shape[0]
    fx, fy = loc[:, :, 1] - (100.0 - R // 2), loc[:, :, 2] - (100.0 - R // 2)
    fv = loc[:, :, 0] > 0
    tv = truth[:, :, 4] > 0
    D = (fx[:, None, :] - truth[:, :, 1][:, :, None]) ** 2 + (fy[:, None, :] - truth[:, :, 2][:, :, None]) ** 2
    D[~(tv[:, :, None] & fv[:, None, :])] = np.inf
    d2, crb, roi, dz = [], [], [], []
    rows = np.arange(n)
    for _ in range(3):
        flat = D.reshape(n, 9)
        k = flat.argmin(1)
        best = flat[rows, k]
        ok = best < 1.0
        ti, fi = k // 3, k % 3
        d2.append(best[ok])
        c = loc[rows, fi, 9:11][ok] / pixel_size
        crb.append((c ** 2).sum(1))   # CRLBx^2 + CRLBy^2: the bound on E[dx^2 + dy^2]
        roi.append(rows[ok])
        dz.append(loc[rows, fi, 3][ok] - truth[rows, ti, 3][ok])
        D[rows[ok], ti[ok], :] = np.inf
        D[rows[ok], :, fi[ok]] = np.inf
    d2, crb, roi, dz = np.concatenate(d2), np.concatenate(crb), np.concatenate(roi), np.concatenate(dz)
    # RMSE against the CRLB over the matched emitters with a finite, informative bound (< 0.5 px per
    # axis); near-degenerate multi-emitter fits have unbounded CRLBs that would swamp the mean
    ok = np.isfinite(crb) & (crb > 0) & (crb < 0.5)
    # the same over the ROIs whose fitted emitter count equals the true count (the model the CRLB
    # is computed for), where the fit's precision, not the count selection, sets the error
    right = (fv.sum(1) == tv.sum(1))[roi] & ok
    zs = {"z_rmse_nm": float(np.sqrt(np.mean(dz ** 2))), "z_median_abs_err_nm": float(np.median(np.abs(dz)))} if zerr else {}
    return {**zs, "rmse_px": float(np.sqrt(d2.mean())), "recall": float(len(d2) / max(tv.sum(), 1)),
            "precision": float(len(d2) / max(fv.sum(), 1)),
            "rmse_over_crlb": float(np.sqrt(d2[ok].mean() / crb[ok].mean())) if ok.any() else None,
            "crlb_informative_fraction": float(ok.mean()) if len(ok) else None,
            "count_correct_fraction": float((fv.sum(1) == tv.sum(1)).mean()),
            "rmse_over_crlb_count_correct": float(np.sqrt(d2[right].mean() / crb[right].mean())) if right.any() else None,
            "rmse_over_crlb_count_correct_by_count": {
                str(k): (float(np.sqrt(d2[sel].mean() / crb[sel].mean())) if sel.any() else None)
                for k in (1, 2, 3) for sel in [right & (tv.sum(1)[roi] == k)]}}


def fp64_error(params, rois, wle, rows, m):
    """Error of single-emitter rows against the FP64 oracle MLE (Levenberg-Marquardt on the
    unweighted Poisson likelihood, run to convergence from the row's parameters) over the first m
    kept rows.  The reference schedule weights all but MLEFit ROIs (WLE), so the two estimators
    share an optimum only for MLEFit ROIs; the classes are reported separately."""
    import oracle_lib as O   # checker only
    R, half = params.ROISize, params.ROISize // 2
    keep = np.nonzero(rows[:, 0] > 0)[0][:m]
    out = {}
    for name, cls in (("mlefit", 0), ("weighted", None)):   # WLE type 0: MLEFit (unweighted)
        idx = keep[(wle[keep, 5] == cls) if cls is not None else (wle[keep, 5] != 0)]
        dxy, dn, db, conv = [], [], [], 0
        for i in idx:
            img = rois[i, :R * R].astype(np.float64).reshape(R, R)
            q = np.maximum((img - params.Offset) * np.float32(params.KAdc), 1.0) / params.QE
            xc, yc = float(rois[i, R * R]), float(rois[i, R * R + 1])
            x, y = rows[i, 1] - xc + half, rows[i, 2] - yc + half
            t, it, _ = O.mle(q, K=1, theta0=np.array([rows[i, 6], x, y, rows[i, 4], rows[i, 7]], np.float64))
            conv += it < 200
            dxy.append(max(abs(t[1] - x), abs(t[2] - y)))
            dn.append(abs(t[0] - rows[i, 6]) / t[0])
            db.append(abs(t[4] - rows[i, 7]) / t[4])
        if not dxy:
            continue
        dxy, dn, db = map(np.array, (dxy, dn, db))
        out[name] = {"rows": int(len(idx)), "oracle_converged": conv / len(idx),
                     "dxy_px_p50": float(np.median(dxy)), "dxy_px_p99": float(np.percentile(dxy, 99)),
                     "photons_rel_p50": float(np.median(dn)), "photons_rel_p99": float(np.percentile(dn, 99)),
                     "bkg_rel_p50": float(np.median(db)), "bkg_rel_p99": float(np.percentile(db, 99))}
    return out


def multi_emitter_leg(n, seed, log):
    """Plain (K = 1..3 converged fits + FP64 selection) vs warm-started multi-emitter fitting of
    n simulator 7x7 ROIs with 1-3 emitters 1-4 px apart; throughputs from library HIP events."""
    import torch
    from qcstorm import warmstart as WS
    p = WS.params()
    rois, truth = WS.sim_dev(n, seed)
    if os.path.exists(WS.DEFAULT_WEIGHTS):
        net, wsrc = WS.load(), "qc-storm_amd/weights/warmstart_r7.npy"
    else:
        log("bench: no saved warm-start weights; training 2^19 ROIs in-run")
        net, _ = WS.train(n_rois=1 << 19, log=log)
        wsrc = "trained in-run on 2^19 ROIs"
    ctx = Q.Context(p)
    loc_p = torch.zeros((3 * n, 12), dtype=torch.float32, device="cuda")
    loc_w = torch.zeros((3 * n, 12), dtype=torch.float32, device="cuda")
    ctx.fit_mle_dev(rois, n, 3, loc_p)
    st = ctx.fit_warm_dev(net, rois, n, loc_w)
    stp = ctx.fit_plain_stats_dev(rois, n, 3, loc_p)
    torch.cuda.synchronize()

    def timed(fn, reps=5):
        """median over reps calls of the library's HIP-event time (fits + network)"""
        ts = []
        for _ in range(reps):
            ctx.timing_enable(True)
            fn()
            t = ctx.timing()
            ts.append(t["cfit_ms"] + t["nn_ms"])
        return float(np.median(ts)), ts

    plain_ms, plain_all = timed(lambda: ctx.fit_mle_dev(rois, n, 3, loc_p))
    warm_ms, warm_all = timed(lambda: ctx.fit_warm_dev(net, rois, n, loc_w, stats=False))
    ctx.timing_enable(True)
    ctx.fit_warm_dev(net, rois, n, loc_w, stats=False)
    tw = ctx.timing()
    torch.cuda.synchronize()
    m = min(n, 65536)
    th = truth[:m].cpu().numpy()
    rp = match_rmse(loc_p[:3 * m].cpu().numpy().reshape(m, 3, 12), th, p.PixelSize)
    rw = match_rmse(loc_w[:3 * m].cpu().numpy().reshape(m, 3, 12), th, p.PixelSize)
    ctx.close()
    net.close()
    # evaluations: device-counted loss-and-gradient evaluations of the plain path (pixel evaluations
    # = evaluations x 49).  Per pixel a K-emitter evaluation costs 23 + 14 (K - 1) FLOP (FMA = 2:
    # the model, the deviance term and the 3K + 2 gradient sums) and 3 transcendentals (exp of the
    # row term aside); the count does not split by K, so the roofline below prices every
    # evaluation at K = 1 and is a lower bound.
    evals = stp["evals"]
    flop_lb = 23.0 * 49 * evals
    ach_lb = flop_lb / (plain_ms * 1e-3) / 1e12
    return {"rois": n, "plain_fits_per_s": n / (plain_ms * 1e-3), "warm_fits_per_s": n / (warm_ms * 1e-3),
            "warm_speedup": plain_ms / warm_ms, "nn_ms": tw["nn_ms"], "plain_ms": plain_ms, "warm_ms": warm_ms,
            "timed_calls": len(plain_all), "plain_ms_all": plain_all, "warm_ms_all": warm_all,
            "plain_evaluations": int(evals), "plain_pixel_evaluations_per_s": evals * 49 / (plain_ms * 1e-3),
            "plain_iterations_mean": stp["iters"] / n,
            "plain_roofline_lower_bound": {"bound": "valu", "achieved": ach_lb, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                                           "frac": ach_lb / PEAK_FP32_TFLOPS, "flop_per_pixel_evaluation": 23,
                                           "pixel_evaluations": int(evals) * 49},
            "plain": rp, "warm": rw, "rmse_ratio_warm_vs_plain": rw["rmse_px"] / rp["rmse_px"],
            "rmse_over_crlb_plain": rp["rmse_over_crlb"], "rmse_over_crlb_warm": rw["rmse_over_crlb"],
            "rmse_over_crlb_count_correct_plain": rp["rmse_over_crlb_count_correct"],
            "rmse_over_crlb_count_correct_warm": rw["rmse_over_crlb_count_correct"],
            "fallback_fraction": st["fallbacks"] / n, "warm_iterations_mean": st["iters"] / n,
            "weights": wsrc, "accuracy_sample_rois": m}


def as3d_multi_emitter_leg(n, seed, log):
    """Converged fitting of the elliptical model (per-emitter widths, K = 1..3, up to 16
    parameters, FP64 selection and CRLB, z from the calibration polynomial) over n simulator 9x9
    astigmatic ROIs with 1-3 emitters 1-4 px apart; throughput from library HIP events."""
    import torch
    R, kmax = 9, 3
    sp = Q.default_sim(width=R, height=R, astig=1, n_per_unit_max=kmax, min_sep=1.0, max_sep=4.0, photons_min=1000.0,
                       photons_max=8000.0, bg_min=20.0, bg_max=80.0)
    rois = torch.empty((n, Q.roi_words(R)), dtype=torch.int16, device="cuda")
    truth = torch.empty((n, kmax, Q.TRUTH_FLOATS), dtype=torch.float32, device="cuda")
    Q.check(Q.lib().qcs_sim_rois_dev(C.byref(sp), seed, n, 1, C.c_void_p(rois.data_ptr()), C.c_void_p(truth.data_ptr()), None),
            "qcs_sim_rois_dev")
    # sigma_x/y = 1.2 sqrt(1 + ((z -+ 250) / 400)^2): z = -111.11 (sigma_x^2 - sigma_y^2)
    p = Q.default_params(QE=1.0, ROISize=R, LocType=Q.LOCTYPE_AS3D, PixelSize=100.0, p1_XGY=-111.11, p1_XLY=-111.11,
                         ZDepthCorrFactor=1.0)
    ctx = Q.Context(p)
    loc = torch.zeros((3 * n, 12), dtype=torch.float32, device="cuda")
    sel = torch.zeros(n, dtype=torch.int32, device="cuda")
    ctx.fit_mle_dev(rois, n, kmax, loc, sel)
    torch.cuda.synchronize()
    ts = []
    for _ in range(5):
        ctx.timing_enable(True)
        ctx.fit_mle_dev(rois, n, kmax, loc, sel)
        ts.append(ctx.timing()["cfit_ms"])
    t = {"cfit_ms": float(np.median(ts))}
    st = ctx.fit_plain_stats_dev(rois, n, kmax, loc, sel)
    torch.cuda.synchronize()
    m = min(n, 65536)
    th = truth[:m].cpu().numpy()
    acc = match_rmse(loc[:3 * m].cpu().numpy().reshape(m, 3, 12), th, p.PixelSize, R=R, zerr=True)
    s_h, k_true = sel[:m].cpu().numpy(), (th[:, :, 4] > 0).sum(1)
    ctx.close()
    return {"rois": n, "roi": R, "fits_per_s": n / (t["cfit_ms"] * 1e-3), "ms": t["cfit_ms"], "ms_all": ts,
            "iterations_mean": st["iters"] / n, "final_counts": [st["final_k1"], st["final_k2"], st["final_k3"]],
            "count_accuracy_by_true_count": {str(k): float((s_h[k_true == k] == k).mean()) for k in (1, 2, 3)},
            "accuracy": acc, "accuracy_sample_rois": m,
            "data": "simulator: photons log-U(1000, 8000), background U(20, 80), z U(-500, 500) nm"}


def hd_frames_leg(nframes, batch, seed, log, warm_rois=1 << 20, reps=5):
    """High-density frames to localizations through the converged multi-emitter path.

    Workload: nframes uint16 frames of 1024 x 1024 px, simulated on the GPU (seed) at 2-5
    emitters / um^2 (0.02-0.05 per 100 nm px, per-frame density uniform in that range), emitters
    in sites of 1-3 members 1-2.5 px apart (pairs and triples closer than 2 sigma), photons
    log-U(500, 5000), background U(10, 100).  Each batch of `batch` frames is one HIP graph
    replay of qcs_loc_mle_frames_dev: mixed-size extraction (7/9/11), converged 1-3 emitter fits
    with FP64 likelihood-ratio selection and FP64 CRLB in persistent work-queue kernels, rows in
    frame order.  The frames are resident in HBM; each replay takes its batch by a device copy
    into the graph's input buffer (part of the timed work).

    The warm-start comparison (the learned network ahead of the 2D fits) runs on the 7 x 7 ROIs
    the same extraction cuts from these frames (up to warm_rois of them), plain vs warm, each
    timed as the median of `reps` calls, with accuracy against the simulated emitters."""
    import torch
    from scipy.spatial import cKDTree
    from qcstorm import warmstart as WS
    W = H = 1024
    sp = Q.default_sim(width=W, height=H, density_per_px=0.02, density_max_per_px=0.05, n_per_unit_max=3, min_sep=1.0,
                       max_sep=2.5, qe=1.0)
    p = Q.default_params(QE=1.0, ImageWidth=W, ImageHigh=H, ROISize=7, MultiEmitterFitEn=1, PixelSize=100.0,
                         BatchedImageNum=1)
    frames = torch.empty((nframes, H, W), dtype=torch.int16, device="cuda")
    for f0 in range(0, nframes, 25):
        nf = min(25, nframes - f0)
        Q.check(Q.lib().qcs_sim_frames_dev(C.byref(sp), seed, 1 + f0, nf, C.c_void_p(frames[f0].data_ptr()), None),
                "qcs_sim_frames_dev")
    torch.cuda.synchronize()
    ctx = Q.Context(p)
    cap = 40000 * batch
    staging = frames[:batch].clone()
    d_out = torch.empty((3 * cap, 12), dtype=torch.float32, device="cuda")
    d_n = torch.zeros(4, dtype=torch.int32, device="cuda")
    d_start = torch.ones(1, dtype=torch.int32, device="cuda")
    g = ctx.graph_frames(staging, batch, d_start, cap, 3, d_out, d_n)
    nb = nframes // batch
    counts = torch.zeros((nb, 4), dtype=torch.int32, device="cuda")
    first_rows = None

    def run_all():
        for b in range(nb):
            staging.copy_(frames[b * batch:(b + 1) * batch])
            d_start.fill_(1 + b * batch)
            g.launch()
            counts[b].copy_(d_n)

    run_all()   # warm-up pass
    torch.cuda.synchronize()
    first_n = int(counts[0, 0].item())
    staging.copy_(frames[:batch])
    d_start.fill_(1)
    g.launch()
    torch.cuda.synchronize()
    first_rows = d_out[:int(d_n[0].item())].cpu().numpy()
    t0 = time.perf_counter()
    run_all()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    cnt = counts.cpu().numpy().astype(np.int64)
    rows_tot, rois_tot, cand_tot = int(cnt[:, 1].sum()), int(cnt[:, 2].sum()), int(cnt[:, 3].sum())
    overflow = bool((cnt[:, 1] > cnt[:, 0]).any() or (cnt[:, 3] > cnt[:, 2]).any())
    # accuracy of the first batch against the simulated emitters (matched within 0.5 px)
    errs, crb, nrows, ntrue, found = [], [], 0, 0, 0
    for f in range(1, batch + 1):
        tr = Q.frame_emitters(sp, seed, f)
        r = first_rows[first_rows[:, 11] == f]
        d, i = cKDTree(tr[:, 1:3]).query(r[:, 1:3]) if len(r) else (np.zeros(0), np.zeros(0, int))
        ok = d < 0.5
        errs.append(d[ok])
        crb.append(np.hypot(r[ok, 9], r[ok, 10]) / p.PixelSize)
        nrows += len(r)
        ntrue += len(tr)
        found += len(np.unique(i[ok]))
    e, c = np.concatenate(errs), np.concatenate(crb)
    fin = np.isfinite(c) & (c > 0) & (c < 0.5)
    g.close()
    res = {"frames": nb * batch, "frame": [W, H], "batch_frames": batch, "graph_per_batch": True,
           "ms": dt * 1e3, "frames_per_s": nb * batch / dt, "rois_per_s": rois_tot / dt, "localizations_per_s": rows_tot / dt,
           "rois": rois_tot, "candidates": cand_tot, "localizations": rows_tot, "capacity_overflow": overflow,
           "first_batch_accuracy": {"emitters": ntrue, "rows": nrows, "recall_0p5px": found / max(ntrue, 1),
                                    "precision_0p5px": float(len(e)) / max(nrows, 1),
                                    "median_error_px": float(np.median(e)) if len(e) else None,
                                    "rmse_px": float(np.sqrt(np.mean(e ** 2))) if len(e) else None,
                                    "rmse_over_crlb": float(np.sqrt(np.mean(e[fin] ** 2) / np.mean(c[fin] ** 2))) if fin.any() else None},
           "data": "simulator seed %d: 2-5 emitters/um^2, sites of 1-3 emitters 1-2.5 px apart, photons log-U(500, 5000), "
                   "background U(10, 100)" % seed}
    log("bench: high-density frames %.0f frames/s, %.3g ROIs/s" % (res["frames_per_s"], res["rois_per_s"]))

    # warm start on the 7 x 7 ROIs of these frames: plain vs warm, median of reps timed calls each
    if warm_rois > 0 and os.path.exists(WS.DEFAULT_WEIGHTS):
        net = WS.load()
        sets7, frames_of = [], []
        got = 0
        for b in range(nb):
            sets = ctx.extract_mixed_dev(frames[b * batch:(b + 1) * batch], batch, 1 + b * batch, cap, (cap, 16, 16))
            R, d_rois, n, _ = sets[0]
            sets7.append(d_rois[:n])
            got += n
            if got >= warm_rois:
                break
        rois7 = torch.cat(sets7)[:warm_rois].contiguous()
        n7 = rois7.shape[0]
        ctx7 = Q.Context(Q.default_params(QE=1.0, ROISize=7, PixelSize=100.0))
        loc_p = torch.zeros((3 * n7, 12), dtype=torch.float32, device="cuda")
        loc_w = torch.zeros_like(loc_p)
        ctx7.fit_mle_dev(rois7, n7, 3, loc_p)
        stw = ctx7.fit_warm_dev(net, rois7, n7, loc_w)
        stp = ctx7.fit_plain_stats_dev(rois7, n7, 3, loc_p)
        torch.cuda.synchronize()

        def med(fn):
            ts = []
            for _ in range(reps):
                ctx7.timing_enable(True)
                fn()
                t = ctx7.timing()
                ts.append(t["cfit_ms"] + t["nn_ms"])
            return float(np.median(ts)), [float(x) for x in ts]

        plain_ms, plain_all = med(lambda: ctx7.fit_mle_dev(rois7, n7, 3, loc_p))
        warm_ms, warm_all = med(lambda: ctx7.fit_warm_dev(net, rois7, n7, loc_w, stats=False))
        torch.cuda.synchronize()
        # accuracy: rows of the ROIs from the first frames against the simulated emitters
        hdr = rois7[:, 49:53].cpu().numpy().view(np.uint16).astype(np.int64)
        fr = hdr[:, 2] + 65536 * hdr[:, 3]
        acc = {}
        for name, loc in (("plain", loc_p), ("warm", loc_w)):
            L = loc.cpu().numpy().reshape(n7, 3, 12)
            e_all, c_all, nr = [], [], 0
            for f in np.unique(fr)[:batch]:
                tr = Q.frame_emitters(sp, seed, int(f))
                r = L[fr == f].reshape(-1, 12)
                r = r[r[:, 0] > 0]
                d, _ = cKDTree(tr[:, 1:3]).query(r[:, 1:3])
                ok = d < 0.5
                e_all.append(d[ok])
                c_all.append(np.hypot(r[ok, 9], r[ok, 10]) / p.PixelSize)
                nr += len(r)
            e2, c2 = np.concatenate(e_all), np.concatenate(c_all)
            fin2 = np.isfinite(c2) & (c2 > 0) & (c2 < 0.5)
            acc[name] = {"rows": nr, "matched_0p5px": int(len(e2)), "rmse_px": float(np.sqrt(np.mean(e2 ** 2))),
                         "rmse_over_crlb": float(np.sqrt(np.mean(e2[fin2] ** 2) / np.mean(c2[fin2] ** 2)))}
        res["warm_start_7x7"] = {
            "rois": n7, "plain_ms_median": plain_ms, "warm_ms_median": warm_ms, "plain_ms_all": plain_all,
            "warm_ms_all": warm_all, "plain_rois_per_s": n7 / (plain_ms * 1e-3), "warm_rois_per_s": n7 / (warm_ms * 1e-3),
            "warm_speedup": plain_ms / warm_ms, "plain_iterations_mean": stp["iters"] / n7,
            "warm_iterations_mean": stw["iters"] / n7, "fallback_fraction": stw["fallbacks"] / n7,
            "accuracy_first_frames": acc,
            "rmse_ratio_warm_vs_plain": acc["warm"]["rmse_px"] / acc["plain"]["rmse_px"]}
        ctx7.close()
        net.close()
    ctx.close()
    return res


def session_leg(frames_dev, params, F, log, group=1):
    """The drop-in session path the plug-ins use (lm_StartLocThread .. lm_StopLocThread): F host
    frames fed through qcs_session_feed in groups of `group` (each copied into pinned memory and
    sent to the device by the session's worker: extraction, fit + CRLB, frame-sorted compaction,
    rendering), rows streamed to a result file by the writer thread; timed from the first feed
    until the session reports finished (file closed).  The frames reach host memory before the
    timed region (a device-to-host copy of the simulated frames)."""
    import tempfile
    h = frames_dev[:F].cpu().numpy().view(np.uint16)
    p = Q.QcsParams.from_buffer_copy(params)
    p.BatchedImageNum = group
    out_dir = tempfile.mkdtemp(prefix="qcs_session_")
    path = os.path.join(out_dir, "bench_result.txt")
    s = Q.Session(p)
    s.set_output(path)
    t0 = time.perf_counter()
    for f0 in range(0, F, group):
        s.feed(h[f0:f0 + group])
    s.finish()
    s.wait(timeout_s=600.0)
    dt = time.perf_counter() - t0
    rows = s.locs().shape[0]
    s.close()
    size = os.path.getsize(path) if os.path.exists(path) else -1
    try:
        os.remove(path)
        os.rmdir(out_dir)
    except OSError:
        pass
    log("bench: session path %.0f frames/s (host frames, H2D and result file included)" % (F / dt))
    return {"frames": F, "feed_group_frames": group, "ms": dt * 1e3, "frames_per_s": F / dt, "rows": int(rows),
            "rows_per_s": rows / dt, "result_file_bytes": int(size), "result_file_complete": size == rows * 48,
            "host_bytes_per_s": F * h.shape[1] * h.shape[2] * 2 / dt}


def as3d_frames_leg(F, stream, log, steps=3):
    """The astigmatic-3D workload from frames: F simulated 1024 x 1024 frames (0.3 emitters/um^2,
    z uniform in +-500 nm, photons 1000-8000, background 20-80), batches of 4 frames, through
    extraction, the 9 x 9 elliptical one-emitter fit + 15 x 15 CRLB, frame-sorted compaction and
    rendering (the same pipeline as the headline step); frames/s over `steps` timed passes."""
    import torch
    W = H = 1024
    params = Q.default_params(ImageWidth=W, ImageHigh=H, ROISize=9, LocType=Q.LOCTYPE_AS3D, MultiEmitterFitEn=0,
                              WLEEn=1, PixelSize=100.0, PixelZoom=10.0, QE=1.0, TotalFrameNum=F, BatchedImageNum=4,
                              p1_XGY=-111.11, p1_XLY=-111.11, ZDepthCorrFactor=1.0)
    Q.lib().qcs_params_update_sr_size(C.byref(params))
    sp = Q.default_sim(width=W, height=H, density_per_px=0.003, astig=1, photons_min=1000.0, photons_max=8000.0,
                       bg_min=20.0, bg_max=80.0, qe=1.0)
    sh = C.c_void_p(stream.cuda_stream)
    frames = torch.empty((F, H, W), dtype=torch.int16, device="cuda")
    with torch.cuda.stream(stream):
        for f0 in range(0, F, 50):
            nf = min(50, F - f0)
            Q.check(Q.lib().qcs_sim_frames_dev(C.byref(sp), 1002, 1 + f0, nf, C.c_void_p(frames[f0].data_ptr()), sh),
                    "qcs_sim_frames_dev")
    stream.synchronize()
    pipe = Pipeline(params, frames, F, 1, stream, torch)
    pipe.step()
    torch.cuda.synchronize()
    pipe.timing_enable()
    t0 = time.perf_counter()
    fits = rows = 0
    for _ in range(steps):
        n, m = pipe.step()
        fits += n
        rows += m
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    tm = pipe.timing()
    pipe.close()
    return {"frames": F, "frame": [W, H], "roi": 9, "batch_frames": 4, "ms_per_pass": dt * 1e3 / steps,
            "frames_per_s": F * steps / dt, "fits_per_s": fits / dt, "rois_per_pass": fits // steps,
            "localizations_per_pass": rows // steps,
            "fit_kernel_fits_per_s": tm["fit_1e_rois"] / (tm["fit_1e_ms"] * 1e-3),
            "extract_ms_per_pass": tm["extract_ms"] / steps, "fit_ms_per_pass": tm["fit_1e_ms"] / steps,
            "crlb_ms_per_pass": tm["crlb_ms"] / steps,
            "data": "simulator seed 1002: 0.3 emitters/um^2, z U(-500, 500) nm, sigma_x/y = 1.2 sqrt(1 + ((z -+ 250)/400)^2), "
                    "photons log-U(1000, 8000), background U(20, 80)"}


class Pipeline:
    """The full-frame pipeline over F resident frames.  With lanes > 1 the frames are split into
    contiguous groups, each extracted, fitted and compacted by its own context on its own stream
    from its own host thread (the GPU runs the lanes' extraction and fitting concurrently); the
    frame-sorted groups are then rendered in order into the SR image."""

    def __init__(self, params, frames, F, lanes, stream, torch):
        import concurrent.futures as cf
        self.torch, self.F, self.stream, self.lanes = torch, F, stream, []
        dev = frames.device
        R = params.ROISize
        bounds = [round(F * k / lanes) for k in range(lanes + 1)]
        for k in range(lanes):
            f0, f1 = bounds[k], bounds[k + 1]
            cap = 20000 * (f1 - f0)
            st = stream if k == 0 else torch.cuda.Stream(device=dev)
            self.lanes.append({
                "f0": f0, "nf": f1 - f0, "cap": cap, "stream": st, "sh": C.c_void_p(st.cuda_stream),
                "ctx": Q.Context(params),
                "rois": torch.empty((cap, Q.roi_words(R)), dtype=torch.int16, device=dev),
                "wle": torch.empty((cap, Q.WLE_FLOATS), dtype=torch.float32, device=dev),
                "loc": torch.empty((cap, Q.REC_FLOATS), dtype=torch.float32, device=dev),
                "out": torch.empty((cap, Q.REC_FLOATS), dtype=torch.float32, device=dev)})
        self.frames = frames
        self.img = torch.empty((params.SRImageHigh, params.SRImageWidth), dtype=torch.float32, device=dev)
        self.ctx = self.lanes[0]["ctx"]
        self.pool = cf.ThreadPoolExecutor(max_workers=lanes) if lanes > 1 else None

    def _lane(self, L):
        n, found = L["ctx"].extract_dev(self.frames[L["f0"]:L["f0"] + L["nf"]], L["nf"], L["f0"] + 1, L["rois"], L["wle"],
                                        L["cap"], L["sh"])
        if found != n:
            raise RuntimeError("ROI buffer too small: %d found, %d kept" % (found, n))
        rows = L["ctx"].fit_dev(L["rois"], L["wle"], n, L["loc"], L["sh"])
        m = L["ctx"].compact_sorted_dev(L["loc"], rows, L["f0"] + 1, L["f0"] + L["nf"], L["out"], L["sh"])
        L["n"], L["m"] = n, m
        return n, m

    def step(self):
        if self.pool is None:
            self._lane(self.lanes[0])
        else:
            list(self.pool.map(self._lane, self.lanes))   # each lane's compaction waits for its stream
        # a fresh SR image per step: the first lane's rows replace the image (clear and draw in one
        # write pass), the other lanes' rows add to it
        for i, L in enumerate(self.lanes):
            draw = self.ctx.render_fresh_dev if i == 0 else self.ctx.render_dev
            draw(L["out"], L["m"], self.img, None, self.lanes[0]["sh"])
        return sum(L["n"] for L in self.lanes), sum(L["m"] for L in self.lanes)

    def timing_enable(self):
        for L in self.lanes:
            L["ctx"].timing_enable(True)

    def timing(self):
        """Stage totals summed over the lanes (each lane's intervals come from its own stream)."""
        tot = None
        for L in self.lanes:
            t = L["ctx"].timing()
            tot = t if tot is None else {k: tot[k] + t[k] for k in tot}
        return tot

    def close(self):
        if self.pool is not None:
            self.pool.shutdown()
        for L in self.lanes:
            L["ctx"].close()


def output_side_legs(pipe, params, dev, sh, W, H, F, torch):
    """Statistics, consecutive-frame merging and drift correction of the step's localizations;
    returns their three JSON blocks."""
    # statistics of the step's localizations (histograms, means, activation density), one call over
    # the frame-sorted rows of lane 0 (all frames when lanes == 1)
    log("bench: statistics leg")
    L0 = pipe.lanes[0]
    st_rows = L0["m"]
    pipe.ctx.loc_stats_dev(L0["out"], st_rows, sh)
    reps = 5
    t0 = time.perf_counter()
    for _ in range(reps):
        st = pipe.ctx.loc_stats_dev(L0["out"], st_rows, sh)
    st_ms = (time.perf_counter() - t0) * 1e3 / reps
    statistics = {"rows": int(st_rows), "ms": st_ms, "rows_per_s": st_rows / (st_ms * 1e-3),
                  "valid": st["valid"], "density_total": st["density_total"],
                  "loc_density_2d_per_um2": st["loc_density_2d"], "mean_total_photon": st["mean_total_photon"],
                  "mean_loc_prec_nm": st["mean_loc_prec"]}

    # consecutive-frame merging of the step's localizations (one call that gives everything back,
    # 100 nm threshold), then removal of the merged-away rows
    log("bench: consecutive-frame merging leg")
    params_c = Q.QcsParams.from_buffer_copy(params)
    params_c.ConsecFitEn = 1
    params_c.ConsecFit_DistanceTh_nm = 100.0
    ctx_c = Q.Context(params_c)
    cons = Q.Consec(ctx_c)
    c_out = torch.empty((st_rows, Q.REC_FLOATS), dtype=torch.float32, device=dev)
    c_kept = torch.empty_like(c_out)
    cons.fit_dev(L0["out"], st_rows, True, c_out, sh)
    reps = 5
    t0 = time.perf_counter()
    for _ in range(reps):
        cm, ratio = cons.fit_dev(L0["out"], st_rows, True, c_out, sh)
    c_ms = (time.perf_counter() - t0) * 1e3 / reps
    ck = ctx_c.compact_dev(c_out, cm, c_kept, sh)
    torch.cuda.synchronize()
    consec = {"rows": int(st_rows), "ms": c_ms, "rows_per_s": st_rows / (c_ms * 1e-3), "threshold_nm": 100.0,
              "rows_after_merge": int(ck), "ontime_ratio": [float(x) for x in ratio]}

    # drift correction by SR-image cross-correlation: a fixed structure (25,000 points, 12,500
    # localizations per frame) drifting over 200 frames of 2048 x 2048 px, groups of 50 frames
    log("bench: drift-correction leg")
    d_rows_h, d_dx, d_dy = drift_workload(W, H, F, 25000, 77)
    params_d = Q.QcsParams.from_buffer_copy(params)
    ctx_d = Q.Context(params_d)
    d_rows = torch.from_numpy(d_rows_h).to(dev)
    corr = min(50, max(F // 4, 1))   # 50-frame groups at the default 200 frames
    ctx_d.drift_correct_dev(d_rows.clone(), d_rows_h.shape[0], corr, stream=sh)   # warm-up (allocates the SR images)
    ctx_d.timing_enable(True)
    reps = 2
    t0 = time.perf_counter()
    for _ in range(reps):
        work = d_rows.clone()
        torch.cuda.synchronize()
        total_d, dfx, dfy, dgxy = ctx_d.drift_correct_dev(work, d_rows_h.shape[0], corr, stream=sh)
    dr_ms = (time.perf_counter() - t0) * 1e3 / reps
    td = ctx_d.timing()
    cc_launches = max(td["drift_cc_launches"], 1)
    cc_ms = td["drift_cc_ms"] / cc_launches
    cc_flop = 2.0 * 47 * 47 * td["drift_cc_pixels"] / cc_launches
    cc_tflops = cc_flop / (cc_ms * 1e-3) / 1e12
    from_frames = [(1 + k * (corr // 2), k * (corr // 2) + corr) for k in range(dgxy.shape[0])]
    mdx = np.array([d_dx[a - 1:b].mean() for a, b in from_frames])
    mdy = np.array([d_dy[a - 1:b].mean() for a, b in from_frames])
    drift = {"rows": int(d_rows_h.shape[0]), "frames": int(total_d), "corr_frames": corr, "groups": int(dgxy.shape[0]),
             "ms": dr_ms, "rows_per_s": d_rows_h.shape[0] / (dr_ms * 1e-3),
             "cc_kernel": {"avg_launch_ms": cc_ms, "launches_per_call": cc_launches // reps,
                           "sr_pixels": td["drift_cc_pixels"] // cc_launches,
                           "roofline": {"bound": "mfma", "achieved": cc_tflops, "peak": 157.3, "unit": "TFLOP/s",
                                        "frac": cc_tflops / 157.3, "flop_per_launch": cc_flop}},
             "group_shift_err_px": float(max(np.abs(dgxy[:, 0] - (mdx - mdx[0])).max(),
                                             np.abs(dgxy[:, 1] - (mdy - mdy[0])).max()))}
    ctx_d.timing_enable(False)
    keep = {"cons": cons, "ctx_c": ctx_c, "ctx_d": ctx_d, "params_c": params_c, "td": td}
    return statistics, consec, drift, keep


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--frames", type=int, default=200)
    ap.add_argument("--math", choices=("exact", "fast"), default="exact")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--multi-rois", type=int, default=1 << 20,
                    help="ROIs of the multi-emitter plain vs warm-start leg (0 skips it)")
    ap.add_argument("--as3d-rois", type=int, default=1 << 18,
                    help="ROIs of the elliptical-model multi-emitter leg (0 skips it)")
    ap.add_argument("--hd-frames", type=int, default=1000,
                    help="1024x1024 high-density frames of the converged frames leg (0 skips it)")
    ap.add_argument("--hd-batch", type=int, default=50, help="frames per HIP-graph batch of the high-density leg")
    ap.add_argument("--as3d-frames", type=int, default=500,
                    help="1024x1024 frames of the astigmatic-3D frames leg (0 skips it)")
    ap.add_argument("--session-frames", type=int, default=200,
                    help="host frames of the drop-in session-path leg (0 skips it)")
    ap.add_argument("--no-fast-leg", action="store_true", help="skip the fast-math pipeline leg")
    ap.add_argument("--lanes", type=int, default=1,
                    help="frame groups processed concurrently (own context, stream and host thread each)")
    ap.add_argument("--headline-only", action="store_true",
                    help="time only the headline step (no other legs, no CPU baseline): for profiling its kernels")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only a plain device copy of known size (HBM counter calibration)")
    ap.add_argument("--leg-only", choices=("multi_emitter", "as3d_multi_emitter"),
                    help="run only this multi-emitter leg and print its block: for profiling its kernels")
    args = ap.parse_args()
    if args.headline_only:
        args.no_fast_leg = args.no_cpu_baseline = True
        args.multi_rois = args.as3d_rois = args.hd_frames = args.session_frames = args.as3d_frames = 0
    if args.gpus != 1:
        print("bench.py measures one GPU only: the localization path runs on a single MI355X "
              "(--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import torch
    if not torch.cuda.is_available():
        raise SystemExit("bench.py needs a GPU (torch.cuda.is_available() is False)")
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    if args.copy_calibration:
        # 1 GiB read + 1 GiB written per launch: the reference point for FETCH_SIZE / WRITE_SIZE
        src = torch.ones(1 << 28, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        for _ in range(max(args.steps, 1)):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read_per_launch": src.numel() * 4, "copy_bytes_written_per_launch": dst.numel() * 4,
                          "launches": max(args.steps, 1)}), flush=True)
        return
    Q.lib()  # raises when libqcstorm.so is missing: there is no fallback
    if args.leg_only:
        if args.leg_only == "multi_emitter":
            out = multi_emitter_leg(args.multi_rois, 1003, log)
        else:
            out = as3d_multi_emitter_leg(args.as3d_rois, 1004, log)
        print(json.dumps({args.leg_only: out}), flush=True)
        return
    stream = torch.cuda.Stream(device=dev)
    sh = C.c_void_p(stream.cuda_stream)

    W = H = 2048
    F = args.frames
    R = 7
    params = Q.default_params(ImageWidth=W, ImageHigh=H, ROISize=R, LocType=Q.LOCTYPE_GS2D, MultiEmitterFitEn=0,
                              WLEEn=1, PixelSize=100.0, PixelZoom=10.0, QE=1.0, TotalFrameNum=F, BatchedImageNum=1,
                              FitMath=Q.MATH_FAST if args.math == "fast" else Q.MATH_EXACT)
    Q.lib().qcs_params_update_sr_size(C.byref(params))
    sp = Q.default_sim(width=W, height=H, density_per_px=0.003, qe=1.0)

    log("bench: simulating %d frames %dx%d on the GPU" % (F, W, H))
    frames = torch.empty((F, H, W), dtype=torch.int16, device=dev)
    with torch.cuda.stream(stream):
        chunk = 25
        for f0 in range(0, F, chunk):
            nf = min(chunk, F - f0)
            Q.check(Q.lib().qcs_sim_frames_dev(C.byref(sp), 1001, 1 + f0, nf,
                                               C.c_void_p(frames[f0].data_ptr()), sh), "qcs_sim_frames_dev")
    stream.synchronize()

    pipe = Pipeline(params, frames, F, args.lanes, stream, torch)
    step = pipe.step

    log("bench: warmup %d step(s)" % args.warmup)
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    pipe.timing_enable()
    log("bench: timing %d step(s)" % args.steps)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    fits = kept = 0
    for _ in range(args.steps):
        n, m = step()
        fits += n
        kept += m
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    tm = pipe.timing()
    log("bench: stage timing %s" % json.dumps(tm))

    value = fits / dt
    fit_launches = max(tm["fit_1e_launches"], 1)
    fit_avg_ms = tm["fit_1e_ms"] / fit_launches
    fits_per_launch = tm["fit_1e_rois"] / fit_launches
    achieved = FLOP_PER_FIT_7X7 * fits_per_launch / (fit_avg_ms * 1e-3) / 1e12
    tpf = traffic_per_fit()
    roofline = {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                "frac": achieved / PEAK_FP32_TFLOPS,
                "traffic": (tpf * fits_per_launch) if tpf is not None else None,
                "kernel": "fit_1e_kernel<GS2D,7>", "flop_per_launch": FLOP_PER_FIT_7X7 * fits_per_launch,
                "avg_launch_ms": fit_avg_ms}

    n_last = fits // max(args.steps, 1)
    stages = {"frames_per_s": F * args.steps / dt,
              "rois_per_step": n_last, "localizations_per_step": kept // max(args.steps, 1),
              "extract_ms_per_step": tm["extract_ms"] / args.steps,
              "fit_ms_per_step": tm["fit_1e_ms"] / args.steps,
              "crlb_ms_per_step": tm["crlb_ms"] / args.steps,
              "compact_ms_per_step": tm["compact_ms"] / args.steps,
              "render_ms_per_step": tm["render_ms"] / args.steps,
              "fit_kernel_fits_per_s": tm["fit_1e_rois"] / (tm["fit_1e_ms"] * 1e-3),
              "extract_frame_GBps": F * W * H * 2 * args.steps / (tm["extract_ms"] * 1e-3) / 1e9}

    # the same pipeline with the transcendental-unit exp/log (the reference's kernels use the
    # hardware approximations too); its rows are compared with the exact ones on the same ROIs
    fast = None
    if args.math == "exact" and not args.no_fast_leg:
        log("bench: fast-math leg (%d step(s))" % args.steps)
        exact_rows = [L["loc"][:L["n"]].clone() for L in pipe.lanes]
        params_f = Q.QcsParams.from_buffer_copy(params)
        params_f.FitMath = Q.MATH_FAST
        pipe_f = Pipeline(params_f, frames, F, args.lanes, stream, torch)
        for _ in range(args.warmup):
            pipe_f.step()
        torch.cuda.synchronize()
        pipe_f.timing_enable()
        t0 = time.perf_counter()
        fits_f = 0
        for _ in range(args.steps):
            fits_f += pipe_f.step()[0]
        torch.cuda.synchronize()
        dt_f = time.perf_counter() - t0
        tf = pipe_f.timing()
        fl = max(tf["fit_1e_launches"], 1)
        f_avg_ms = tf["fit_1e_ms"] / fl
        f_achieved = FLOP_PER_FIT_7X7 * (tf["fit_1e_rois"] / fl) / (f_avg_ms * 1e-3) / 1e12
        a = np.concatenate([r.cpu().numpy() for r in exact_rows])
        b = np.concatenate([L["loc"][:L["n"]].cpu().numpy() for L in pipe_f.lanes])
        both = (a[:, 0] > 0) & (b[:, 0] > 0)
        d = np.maximum(np.abs(a[both, 1] - b[both, 1]), np.abs(a[both, 2] - b[both, 2]))
        fast = {"value": fits_f / dt_f, "unit": "fits/s", "ms_per_step": dt_f * 1e3 / args.steps,
                "frames_per_s": F * args.steps / dt_f,
                "fit_kernel_fits_per_s": tf["fit_1e_rois"] / (tf["fit_1e_ms"] * 1e-3),
                "roofline": {"achieved": f_achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                             "frac": f_achieved / PEAK_FP32_TFLOPS, "avg_launch_ms": f_avg_ms},
                "vs_exact_rows": {"rows": int(n_last), "acceptance_agree": float(np.mean((a[:, 0] > 0) == (b[:, 0] > 0))),
                                  "dxy_px_median": float(np.median(d)) if d.size else 0.0,
                                  "dxy_px_p99": float(np.percentile(d, 99)) if d.size else 0.0}}
        pipe_f.close()

    statistics = consec = drift = None
    keep = {}   # the output-side legs' contexts, used again by the CPU-baseline checks
    if not args.headline_only:
        statistics, consec, drift, keep = output_side_legs(pipe, params, dev, sh, W, H, F, torch)

    multi = None
    if args.multi_rois > 0:
        log("bench: multi-emitter leg (plain vs warm start, %d ROIs)" % args.multi_rois)
        multi = multi_emitter_leg(args.multi_rois, 1003, log)
    as3d = None
    if args.as3d_rois > 0:
        log("bench: elliptical-model multi-emitter leg (%d 9x9 ROIs)" % args.as3d_rois)
        as3d = as3d_multi_emitter_leg(args.as3d_rois, 1004, log)

    sess = None
    if args.session_frames > 0:
        log("bench: drop-in session path (%d host frames)" % args.session_frames)
        sess = session_leg(frames, params, min(args.session_frames, F), log)

    c0 = None
    if not args.no_cpu_baseline:
        log("bench: 100k-ROI single-emitter case (oracle on the host cores vs the library)")
        c0 = c0_leg(log)

    hd = None
    if args.hd_frames > 0:
        log("bench: high-density frames leg (%d frames 1024x1024, graphs of %d frames)" % (args.hd_frames, args.hd_batch))
        hd = hd_frames_leg(args.hd_frames, args.hd_batch, 1003, log)

    as3d_frames = None
    if args.as3d_frames > 0:
        log("bench: astigmatic-3D frames leg (%d frames 1024x1024, 9x9 ROIs)" % args.as3d_frames)
        as3d_frames = as3d_frames_leg(args.as3d_frames, stream, log)
        log("bench: astigmatic-3D frames %.0f frames/s, %.3g fits/s" % (as3d_frames["frames_per_s"], as3d_frames["fits_per_s"]))

    cpu = None
    parity = None
    if not args.no_cpu_baseline:
        log("bench: CPU baseline (oracle, bounded sample)")
        L0 = pipe.lanes[0]   # its ROIs are the first frames' (raster order): the workload's prefix
        h_rois = L0["rois"][:L0["n"]].cpu().numpy().view(np.uint16)
        h_wle = L0["wle"][:L0["n"]].cpu().numpy()
        cpu, ref = cpu_baseline(params, h_rois, h_wle, args.cpu_seconds)
        if args.math == "exact":
            got = L0["loc"][:ref.shape[0]].cpu().numpy()
            same = int(np.all(got.view(np.uint32) == ref.view(np.uint32), axis=1).sum())
            parity = {"rows": int(ref.shape[0]), "bit_exact_rows": same}
            parity["vs_fp64_mle"] = fp64_error(params, h_rois, h_wle, got, 16384)
        # statistics: the oracle (per-frame all-pairs neighbour scan, one thread) on the rows of the
        # first 6 frames, against the device call on the same rows
        import oracle_lib as O   # checker only
        rows_h = L0["out"][:L0["m"]].cpu().numpy()
        sub = rows_h[rows_h[:, 11] <= 6]
        t0 = time.perf_counter()
        st_ref = O.loc_stats(params, sub)
        cpu_ms = (time.perf_counter() - t0) * 1e3
        d_sub = torch.from_numpy(np.ascontiguousarray(sub)).to(dev)
        st_dev = pipe.ctx.loc_stats_dev(d_sub, sub.shape[0], sh)
        statistics["cpu_baseline"] = {"value": sub.shape[0] / (cpu_ms * 1e-3), "unit": "rows/s", "cores": 1,
                                      "kind": "port", "sample": "rows of frames 1-6 (%d rows)" % sub.shape[0]}
        statistics["parity_sample"] = {
            "rows": int(sub.shape[0]),
            "counts_equal": all(st_dev[k] == st_ref[k] for k in ("valid", "density_total", "density_n0", "density_n1")),
            "histograms_equal": all(np.array_equal(st_dev[k], st_ref[k]) for k in
                                    ("hist_total_photon", "hist_loc_prec", "hist_peak_snr", "hist_psf_width"))}
        # consecutive-frame merging: the oracle (the reference's per-row scan, one thread) on the
        # rows of the first 4 frames, against the device call on the same rows
        sub = rows_h[rows_h[:, 11] <= 4]
        ref_c = O.Consec(keep["params_c"])
        t0 = time.perf_counter()
        want, _ = ref_c.fit(sub, True)
        cpu_ms = (time.perf_counter() - t0) * 1e3
        d_sub = torch.from_numpy(np.ascontiguousarray(sub)).to(dev)
        d_o = torch.empty_like(d_sub)
        gm, _ = keep["cons"].fit_dev(d_sub, sub.shape[0], True, d_o, sh)
        got = d_o[:gm].cpu().numpy()
        nan = np.isnan(got)
        consec["cpu_baseline"] = {"value": sub.shape[0] / (cpu_ms * 1e-3), "unit": "rows/s", "cores": 1, "kind": "port",
                                  "sample": "rows of frames 1-4 (%d rows)" % sub.shape[0]}
        consec["parity_sample"] = {"rows": int(sub.shape[0]), "equal": bool(
            got.shape == want.shape and np.array_equal(nan, np.isnan(want)) and
            np.array_equal(got.view(np.uint32)[~nan], want.view(np.uint32)[~nan]))}
        # drift correction: the oracle (the reference's per-shift passes, one thread) on a 96 x 96 px
        # crop of the drifting structure, 60 frames in groups of 20 (4 correlations of 768^2 SR px)
        p_crop = Q.QcsParams.from_buffer_copy(params)
        p_crop.ImageWidth = p_crop.ImageHigh = 96
        crop, _, _ = drift_workload(96, 96, 60, 1200, 78)
        t0 = time.perf_counter()
        O.drift_correct(p_crop, crop, 20)
        cpu_ms = (time.perf_counter() - t0) * 1e3
        drift["cpu_baseline"] = {"value": 4 * 768 * 768 * 47 * 47 / (cpu_ms * 1e-3), "unit": "SR pixel-shift products/s",
                                 "cores": 1, "kind": "port",
                                 "sample": "96x96 px crop, 60 frames, 4 correlations of 768x768 SR px (%.1f s)" % (cpu_ms / 1e3)}
        td = keep["td"]
        drift["pixel_shift_products_per_s"] = td["drift_cc_pixels"] * 47 * 47 / (td["drift_cc_ms"] * 1e-3)
        d_crop = torch.from_numpy(crop.copy()).to(dev)
        ctx_crop = Q.Context(p_crop)
        _, _, _, gxy_crop = ctx_crop.drift_correct_dev(d_crop, crop.shape[0], 20, stream=sh)
        _, _, _, _, rgx, rgy = O.drift_correct(p_crop, crop, 20)
        drift["parity_sample"] = {"groups": int(rgx.shape[0]),
                                  "max_group_shift_diff_px": float(max(np.abs(gxy_crop[:, 0] - rgx).max(),
                                                                       np.abs(gxy_crop[:, 1] - rgy).max()))}
        ctx_crop.close()

    line = {"metric": METRIC, "value": value, "unit": "fits/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": dt * 1e3 / args.steps, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
            "config": {"workload": "full-frame 2D pipeline: %d frames 2048x2048 u16, 0.3 emitters/um^2, "
                                   "extraction + 7x7 single-emitter BFGS fit + CRLB + sorted compaction + "
                                   "10 nm rendering" % F,
                       "frames": F, "frame": [W, H], "roi": R, "fit_math": args.math,
                       "rois_per_step": n_last, "lanes": args.lanes},
            "roofline": roofline, "cpu_baseline": cpu, "stages": stages, "parity_sample": parity,
            "fast_math": fast, "multi_emitter": multi, "as3d_multi_emitter": as3d, "hd_frames": hd, "rois_100k": c0,
            "session_path": sess, "as3d_frames": as3d_frames,
            "statistics": statistics,
            "consecutive_merge": consec,
            "drift_correction": drift}
    for k in ("cons", "ctx_c", "ctx_d"):
        if k in keep:
            keep[k].close()
    pipe.close()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
