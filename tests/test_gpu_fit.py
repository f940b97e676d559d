"""GPU parity of the BFGS fitters against the CPU oracle (bit-exact in EXACT math mode)."""
import numpy as np
import pytest

import oracle_lib as O
import qcstorm as Q

pytestmark = pytest.mark.gpu


def _config0(n, seed=1000, R=7):
    sp = Q.default_sim(width=R, height=R)
    rois, truth = Q.sim_rois_host(sp, seed, n)
    return rois, truth


@pytest.mark.parametrize("R", [7, 5, 9, 13, 15, 17])
def test_gs2d_single_bit_exact(gpu, R):
    rois, _ = _config0(4096, seed=1000 + R, R=R)
    params = Q.default_params(QE=1.0, ROISize=R)
    w = O.wle(rois, R)
    ref, st = O.fit(params, rois, w)
    ctx = Q.Context(params)
    got = ctx.fit_rois(rois, w)
    assert got.shape == ref.shape
    diff = np.argwhere(got.view(np.uint32) != ref.view(np.uint32))
    print("R", R, "rows", ref.shape[0], "accepted", int((ref[:, 0] > 0).sum()), "mismatching words", len(diff))
    if len(diff):
        r = diff[0][0]
        print("first mismatch row", r, "\n gpu", got[r], "\n orc", ref[r])
    assert len(diff) == 0


def test_gs2d_fast_math_statistics(gpu):
    rois, truth = _config0(4096, seed=1007)
    params = Q.default_params(QE=1.0)
    w = O.wle(rois, 7)
    ref, _ = O.fit(params, rois, w)
    params_f = Q.default_params(QE=1.0, FitMath=Q.MATH_FAST)
    got = Q.Context(params_f).fit_rois(rois, w)
    both = (ref[:, 0] > 0) & (got[:, 0] > 0)
    d = np.maximum(np.abs(got[both, 1] - ref[both, 1]), np.abs(got[both, 2] - ref[both, 2]))
    print("fast vs exact: median %.2e p99 %.2e" % (np.median(d), np.percentile(d, 99)))
    # FAST differs from EXACT by FP32 rounding only: same acceptance for >= 99% of ROIs and a
    # median deviation well inside the FP32 noise band of the schedule (FP32 vs FP64 median ~2e-4 px)
    assert np.mean((ref[:, 0] > 0) == (got[:, 0] > 0)) > 0.99
    assert np.median(d) < 2e-3


def _multi_rois(R, n, seed, astig=False):
    sp = Q.default_sim(width=R, height=R, n_per_unit_max=3, min_sep=1.0, max_sep=3.0, astig=int(astig),
                       photons_min=800, photons_max=6000)
    rois, truth = Q.sim_rois_host(sp, seed, n)
    rng = np.random.default_rng(seed)
    rois[:, R * R + 4] = (rng.random(n) < 0.3).astype(np.uint16)    # detector "multi" flag on some ROIs
    return rois, truth


@pytest.mark.parametrize("R,lt", [(7, 0), (9, 0), (11, 0), (7, 1), (9, 1), (5, 0), (13, 0), (5, 1), (13, 1)])
def test_multi_emitter_cascade_bit_exact(gpu, R, lt):
    rois, _ = _multi_rois(R, 3000, 500 + R + 10 * lt, astig=(lt == 1))
    params = Q.default_params(QE=1.0, ROISize=R, LocType=lt, MultiEmitterFitEn=1)
    w = O.wle(rois, R, lt)
    ref, st = O.fit(params, rois, w)
    ctx = Q.Context(params)
    got = ctx.fit_rois(rois, w)
    gst = ctx.stats()
    print("R", R, "lt", lt, "oracle", st, "gpu", {k: gst[k] for k in ("n_single", "n_2e", "n_3e", "n_rejected", "n_rows")})
    assert (gst["n_single"], gst["n_2e"], gst["n_3e"], gst["n_rejected"], gst["n_rows"]) == \
        (st["n_single"], st["n_2e"], st["n_3e"], st["n_rejected"], st["n_rows"])
    assert got.shape == ref.shape
    bad = np.argwhere(got.view(np.uint32) != ref.view(np.uint32))
    if len(bad):
        r = bad[0][0]
        print("first mismatch row", r, "\n gpu", got[r], "\n orc", ref[r])
    assert len(bad) == 0


@pytest.mark.parametrize("R", [7, 9, 13, 15, 17])
def test_as3d_single_bit_exact(gpu, R):
    sp = Q.default_sim(width=R, height=R, astig=1, photons_min=1000, photons_max=8000, bg_min=20, bg_max=80)
    rois, _ = Q.sim_rois_host(sp, 900 + R, 4096)
    params = Q.default_params(QE=1.0, ROISize=R, LocType=1, p1_XGY=-111.11, p1_XLY=-111.11)
    w = O.wle(rois, R, 1)
    ref, _ = O.fit(params, rois, w)
    got = Q.Context(params).fit_rois(rois, w)
    bad = np.argwhere(got.view(np.uint32) != ref.view(np.uint32))
    print("R", R, "accepted", int((ref[:, 0] > 0).sum()), "mismatch", len(bad))
    assert len(bad) == 0


def test_multi_emitter_rejects_sizes_above_13(gpu):
    """The reference has no 2- and 3-emitter kernels for ROI sizes 15 and 17
    (bfgsMLE_Multi_Emitter_Fit.cu:44-75); the library refuses them instead of skipping stages."""
    R = 15
    sp = Q.default_sim(width=R, height=R)
    rois, _ = Q.sim_rois_host(sp, 5, 8)
    params = Q.default_params(QE=1.0, ROISize=R, MultiEmitterFitEn=1)
    with pytest.raises(Q.QcsError) as e:
        Q.Context(params).fit_rois(rois, O.wle(rois, R))
    assert e.value.rc == Q.QCS_ERR_ARG
