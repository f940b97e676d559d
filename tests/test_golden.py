"""Golden fit vectors (tests/golden/golden_fits.npz, written by tests/golden/make_golden.py):
the reference's MATLAB test ROIs plus seeded simulator ROIs, with the oracle's fit + CRLB rows
for 2D / AS3D, single- and multi-emitter modes.  The CPU test re-runs the oracle on the stored
inputs; the GPU test runs the HIP fitters through the C-ABI.  Both must match bit for bit."""
import os

import numpy as np
import pytest

import oracle_lib as O
import qcstorm as Q

GOLD = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "golden_fits.npz"))
MODES = sorted({k.split("__")[0] for k in GOLD.files})


def _params(name):
    R, lt, mu = (int(v) for v in GOLD[name + "__mode"])
    return Q.default_params(QE=1.0, ROISize=R, LocType=lt, MultiEmitterFitEn=mu, p1_XGY=-111.11, p1_XLY=-111.11,
                            ZDepthCorrFactor=1.0)


@pytest.mark.parametrize("name", MODES)
def test_oracle_reproduces_golden(name):
    rois, wle, loc = GOLD[name + "__rois"], GOLD[name + "__wle"], GOLD[name + "__loc"]
    R, lt, _ = (int(v) for v in GOLD[name + "__mode"])
    assert np.array_equal(O.wle(rois, R, lt).view(np.uint32), wle.view(np.uint32))
    got, st = O.fit(_params(name), rois, wle)
    assert np.array_equal(got.view(np.uint32), loc.view(np.uint32))
    assert [st[k] for k in ("n_in", "n_single", "n_2e", "n_3e", "n_rejected", "n_rows")] == list(GOLD[name + "__stats"])


@pytest.mark.gpu
@pytest.mark.parametrize("name", MODES)
def test_gpu_matches_golden(gpu, name):
    rois, wle, loc = GOLD[name + "__rois"], GOLD[name + "__wle"], GOLD[name + "__loc"]
    ctx = Q.Context(_params(name))
    got = ctx.fit_rois(rois, wle)
    st = ctx.stats()
    ctx.close()
    print(name, "rows", got.shape[0], loc.shape[0])
    assert got.shape == loc.shape
    assert np.array_equal(got.view(np.uint32), loc.view(np.uint32))
    assert [st[k] for k in ("n_in", "n_single", "n_2e", "n_3e", "n_rejected", "n_rows")] == list(GOLD[name + "__stats"])
