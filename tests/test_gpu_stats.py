"""GPU statistics of localization rows (qcs_loc_stats_dev) vs the oracle: histograms, valid and
neighbour counts exact; FP64 sums to 1e-12 (summation order differs); means and the density to
float rounding of those sums."""
import numpy as np
import pytest
import torch

import oracle_lib as O
import qcstorm as Q

pytestmark = pytest.mark.gpu

INT_KEYS = ("valid", "density_first_frame", "density_last_frame", "density_total", "density_n0", "density_n1")
HIST_KEYS = ("hist_total_photon", "hist_loc_prec", "hist_peak_snr", "hist_psf_width")
SUM_KEYS = ("sum_loc_prec", "sum_peak_snr", "sum_psf_width", "sum_background")
MEAN_KEYS = ("mean_total_photon", "mean_loc_prec", "mean_peak_snr", "mean_background", "mean_psf_width",
             "mean_psf_width_ctl", "loc_density_2d")


def _rows(n, seed, nframes, W, H, dense=False):
    rng = np.random.default_rng(seed)
    loc = np.zeros((n, 12), np.float32)
    loc[:, 0] = rng.uniform(0, 500, n)
    loc[:, 1] = rng.uniform(-2, W + 2, n)
    loc[:, 2] = rng.uniform(-2, H + 2, n)
    if dense:   # clusters: many molecules with one or more neighbours
        nc = max(1, n // 4)
        c = rng.integers(0, nc, n)
        loc[:, 1] = np.float32(rng.uniform(0, W, nc)[c] + rng.normal(0, 6, n))
        loc[:, 2] = np.float32(rng.uniform(0, H, nc)[c] + rng.normal(0, 6, n))
    loc[:, 4] = rng.uniform(0.0, 2.5, n)
    loc[:, 5] = rng.uniform(0.0, 2.5, n)
    loc[:, 6] = rng.choice([20.0, 1e3, 5e3, 2e4, 1e12, np.inf], n) * rng.uniform(0.5, 1.5, n)
    loc[:, 7] = rng.uniform(0, 60, n)
    loc[:, 8] = rng.choice([1.0, 5.0, 30.0, 80.0, 1e9], n) * rng.uniform(0.5, 1.5, n)
    loc[:, 9] = rng.choice([np.nan, -1.0, 0.005, 3.0, 12.0, 30.0, 45.0], n) * rng.uniform(0.5, 1.5, n)
    loc[:, 10] = rng.uniform(0, 40, n)
    loc[:, 11] = np.sort(rng.integers(1, nframes + 1, n))
    return loc


def _check(params, loc):
    ref = O.loc_stats(params, loc)
    ctx = Q.Context(params)
    d = torch.from_numpy(np.ascontiguousarray(loc)).cuda()
    got = ctx.loc_stats_dev(d, loc.shape[0])
    for k in INT_KEYS:
        assert got[k] == ref[k], (k, got[k], ref[k])
    for k in HIST_KEYS:
        np.testing.assert_array_equal(got[k], ref[k], err_msg=k)
    for k in SUM_KEYS:
        assert got[k] == pytest.approx(ref[k], rel=1e-12, abs=1e-300), k
    for k in MEAN_KEYS:
        assert got[k] == pytest.approx(ref[k], rel=1e-6, abs=1e-30), k
    return got


@pytest.mark.parametrize("lt,multi,dense", [(0, 0, False), (0, 1, True), (1, 0, True)])
def test_stats_match_oracle(gpu, lt, multi, dense):
    W, H = 512, 384
    loc = _rows(60000, 10 + lt + 2 * multi, 40, W, H, dense)
    params = Q.default_params(ImageWidth=W, ImageHigh=H, LocType=lt, MultiEmitterFitEn=multi)
    got = _check(params, loc)
    assert got["valid"] > 1000 and got["density_total"] > 1000
    assert got["density_n0"] > 0 and got["density_n1"] > 0 and got["loc_density_2d"] > 0


@pytest.mark.parametrize("frames", [(5,), (5, 6), (5, 6, 7), (5, 6, 7, 8)])
def test_stats_short_frame_ranges(gpu, frames):
    loc = _rows(4000, 20 + len(frames), 1, 256, 256, dense=True)
    loc[:, 11] = np.sort(np.random.default_rng(3).choice(frames, loc.shape[0]))
    loc[0, 11] = frames[0]
    loc[-1, 11] = frames[-1]
    _check(Q.default_params(ImageWidth=256, ImageHigh=256), loc)


def test_stats_frame_quirks_and_tiny_batches(gpu):
    p = Q.default_params(ImageWidth=128, ImageHigh=128)
    # zero frames at both ends: the first/last non-zero frames bound the range; row 0 is never
    # the last-frame row
    loc = _rows(3000, 30, 12, 128, 128)
    loc[:40, 11] = 0
    loc[-25:, 11] = 0
    _check(p, loc)
    for n in (1, 2, 3, 50, 51):
        _check(p, _rows(n, 31 + n, 2, 128, 128, dense=True))
    one = _rows(1, 40, 1, 128, 128)
    one[0, 11] = 9
    got = _check(p, one)
    assert (got["density_first_frame"], got["density_last_frame"]) == (9, 0)
    empty = Q.Context(p).loc_stats_dev(torch.zeros((1, 12), device="cuda"), 0)
    assert empty["valid"] == 0 and empty["density_total"] == 0


def test_stats_of_fitted_batch(gpu):
    """Statistics over the rows the fitting pipeline writes (frame-sorted compaction of a
    simulated batch)."""
    sp = Q.default_sim()
    rois, _ = Q.sim_rois_host(sp, 7, 20000)
    params = Q.default_params()
    ctx = Q.Context(params)
    d_rois = torch.from_numpy(rois).cuda()
    wle = torch.zeros((rois.shape[0], Q.WLE_FLOATS), dtype=torch.float32, device="cuda")
    ctx.wle_dev(d_rois, rois.shape[0], wle)
    loc = torch.zeros((rois.shape[0], 12), dtype=torch.float32, device="cuda")
    rows = ctx.fit_dev(d_rois, wle, rois.shape[0], loc)
    torch.cuda.synchronize()
    host = loc[:rows].cpu().numpy()
    host[:, 11] = np.sort(np.random.default_rng(5).integers(1, 30, host.shape[0]))
    got = _check(params, host)
    assert got["valid"] > 0.5 * host.shape[0]
