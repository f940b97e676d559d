"""Tolerance of the GPU fitter against the reference's own arithmetic order, on the 100k seeded
7 x 7 ROIs of the single-emitter reference case (seed 1000).

The GPU (EXACT mode) is bit-exact with the oracle's default arithmetic: the separable Gaussian
exp(-s dx^2) exp(-s dy^2) and the library's deterministic exp/log.  The reference evaluates one
exponential per pixel with the CUDA intrinsics (bfgsMLE_GS2D_core.h:497-503); the oracle's literal
mode (MODEL_LITERAL) restates that order with libm exp/log (TMODE_LIBM), the closest reproducible
stand-in for __expf/__logf.  The fixed BFGS schedule amplifies FP32 rounding differences, so the
two agree to |dx| <= 1e-3 px for about 86% of the accepted fits, not all of them.  Against the
same schedule run in FP64 the default arithmetic is the closer of the two (94% vs 86% within
1e-3 px): the reference order's own FP32 rounding is what separates them, and it bounds what any
FP32 implementation of the reference (its own __expf/__logf included) can claim.  The asserted
percentiles:

    quantity              p50        p90        p99       share within 1e-3
    |dx|, |dy| (px)       <= 3e-4    <= 4e-3    <= 5e-2   >= 0.80
    photons, bg (rel.)    <= 5e-4    <= 5e-3    <= 8e-2   >= 0.78
    acceptance agreement  >= 0.99
"""
import numpy as np
import pytest

import oracle_lib as O
import qcstorm as Q


def _c0(n=100000):
    sp = Q.default_sim()
    rois, _ = Q.sim_rois_host(sp, 1000, n)
    p = Q.default_params(QE=1.0)
    return p, rois, O.wle(rois, 7)


def _compare(a, b):
    both = (a[:, 0] > 0) & (b[:, 0] > 0)
    out = {"agree": float(np.mean((a[:, 0] > 0) == (b[:, 0] > 0)))}
    for name, col, rel in (("dx", 1, False), ("dy", 2, False), ("photons", 0, True), ("bg", 7, True)):
        d = np.abs(a[both, col] - b[both, col])
        if rel:
            d = d / np.abs(b[both, col])
        out[name] = (np.percentile(d, [50, 90, 99]), float(np.mean(d <= 1e-3)))
    return out


def _check(t):
    assert t["agree"] >= 0.99
    for name in ("dx", "dy"):
        (p50, p90, p99), within = t[name]
        assert p50 <= 3e-4 and p90 <= 4e-3 and p99 <= 5e-2 and within >= 0.80, (name, t[name])
    for name in ("photons", "bg"):
        (p50, p90, p99), within = t[name]
        assert p50 <= 5e-4 and p90 <= 5e-3 and p99 <= 8e-2 and within >= 0.78, (name, t[name])


def _report(label, t):
    print(label, "acceptance agreement %.4f" % t["agree"])
    for k in ("dx", "dy", "photons", "bg"):
        (p50, p90, p99), within = t[k]
        print("  %-8s p50 %.2e p90 %.2e p99 %.2e  within 1e-3: %.4f" % (k, p50, p90, p99, within))


def test_oracle_default_vs_reference_order_arithmetic():
    """CPU: the oracle's default arithmetic (what the GPU reproduces bit for bit) against the
    reference-order arithmetic, and both against the FP64 schedule."""
    p, rois, w = _c0()
    a = O.fit_omp(p, rois, w, 8)
    with O.modes(O.TMODE_LIBM, O.MODEL_LITERAL):
        b = O.fit_omp(p, rois, w, 8)
    t = _compare(a, b)
    _report("default vs literal+libm", t)
    _check(t)
    # the same spread separates either FP32 arithmetic from the schedule run in FP64
    m = 20000
    c, _ = O.fit_f64(p, rois[:m], w[:m])
    ta, tb = _compare(a[:m], c[:m].astype(np.float32)), _compare(b[:m], c[:m].astype(np.float32))
    _report("default vs FP64 schedule", ta)
    _report("literal+libm vs FP64 schedule", tb)
    # the default arithmetic is at least as close to FP64 as the reference order is
    assert ta["dx"][1] > tb["dx"][1] - 0.02 and t["dx"][0][1] < 1.5 * tb["dx"][0][1]


@pytest.mark.gpu
def test_gpu_vs_reference_order_arithmetic(gpu):
    p, rois, w = _c0()
    got = Q.Context(p).fit_rois(rois, w)
    ref = O.fit_omp(p, rois, w, 16)
    same = np.all(got.view(np.uint32) == ref.view(np.uint32), axis=1)
    assert same.all(), "GPU rows differ from the oracle's default arithmetic: %d of %d" % ((~same).sum(), len(same))
    with O.modes(O.TMODE_LIBM, O.MODEL_LITERAL):
        lit = O.fit_omp(p, rois, w, 16)
    t = _compare(got, lit)
    _report("GPU vs literal+libm", t)
    _check(t)
