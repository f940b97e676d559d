"""Write the golden fit vectors: ROI records (the reference's MATLAB test ROIs plus seeded
simulator ROIs) and the oracle's fit + CRLB rows for each mode, in golden_fits.npz.

The reference ships input ROIs only (no known answers), so these vectors come from the oracle.
They freeze its output: the CPU tests re-run the oracle against them and the GPU tests compare
the HIP fitters with them.  Regenerate only on a deliberate change of the oracle.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import oracle_lib as O  # noqa: E402
import qcstorm as Q  # noqa: E402

# mode name -> (ROISize, LocType, MultiEmitterFitEn, MAT ROI key prefixes, simulator overrides)
MODES = {
    "gs2d_r7": (7, 0, 0, ("mlebfgs_2d__", "mlewt_2d__roi1"), dict()),
    "gs2d_r7_multi": (7, 0, 1, ("multi_2e_2d__", "multi_3e_2d__", "mlebfgs_2d__"),
                      dict(n_per_unit_max=3, min_sep=1.0, max_sep=3.0, photons_min=800.0)),
    "gs2d_r11_multi": (11, 0, 1, ("mlewt_2d__roi2",), dict(n_per_unit_max=2, min_sep=1.5, max_sep=4.0)),
    "as3d_r9": (9, 1, 0, (), dict(astig=1, photons_min=1000.0, photons_max=8000.0, bg_min=20.0, bg_max=80.0)),
    "as3d_r11": (11, 1, 0, ("mlebfgs_as3d__", "mlewt_as3d__", "tubulin_3d__Tubulin_A647_3DRawSel200f_2"), dict(astig=1)),
    "as3d_r7_multi": (7, 1, 1, ("multi_2e_3d__",), dict(astig=1, n_per_unit_max=2, min_sep=1.5, max_sep=3.0)),
}
N_SIM = 256


def roi_records(imgs, first_frame):
    R = imgs.shape[1]
    out = np.zeros((imgs.shape[0], Q.roi_words(R)), np.uint16)
    out[:, :R * R] = imgs.reshape(imgs.shape[0], -1)
    out[:, R * R + 0] = 100
    out[:, R * R + 1] = 100
    fr = first_frame + np.arange(imgs.shape[0])
    out[:, R * R + 2] = fr % 65536
    out[:, R * R + 3] = fr // 65536
    return out


def params(R, lt, mu):
    return Q.default_params(QE=1.0, ROISize=R, LocType=lt, MultiEmitterFitEn=mu, p1_XGY=-111.11, p1_XLY=-111.11,
                            ZDepthCorrFactor=1.0)


def build():
    mat = np.load(os.path.join(HERE, "mat_rois.npz"))
    out = {}
    for name, (R, lt, mu, prefixes, simkw) in MODES.items():
        imgs = [mat[k] for k in sorted(mat.files) if k.startswith(prefixes) and mat[k].shape == (R, R)]
        parts = []
        if imgs:
            parts.append(roi_records(np.stack(imgs), 1))
        sp = Q.default_sim(width=R, height=R, **simkw)
        sim, _ = Q.sim_rois_host(sp, 4000 + R + 10 * lt + 100 * mu, N_SIM, first_frame=1 + len(imgs))
        parts.append(sim)
        rois = np.concatenate(parts)
        p = params(R, lt, mu)
        wle = O.wle(rois, R, lt)
        loc, st = O.fit(p, rois, wle)
        out[name + "__rois"] = rois
        out[name + "__wle"] = wle
        out[name + "__loc"] = loc
        out[name + "__stats"] = np.array([st[k] for k in ("n_in", "n_single", "n_2e", "n_3e", "n_rejected", "n_rows")],
                                         np.int32)
        out[name + "__mode"] = np.array([R, lt, mu], np.int32)
    return out


if __name__ == "__main__":
    data = build()
    np.savez_compressed(os.path.join(HERE, "golden_fits.npz"), **data)
    print({k: v.shape for k, v in data.items() if k.endswith("__loc")})
