"""Extract the reference's MATLAB test inputs (.mat files, read as data) into small npz fixtures.

Run in a container that holds the reference tree; the fixtures it writes are committed so the
tests never read the reference at run time.  Only numeric arrays are copied (ROI images and the
spline-test vectors); no reference code is read or run.
"""
import os
import sys

import numpy as np
import scipy.io as sio

REF = "/root/reference/MATLAB codes"
OUT = os.path.dirname(os.path.abspath(__file__))

SOURCES = {
    "mlebfgs_2d": "MLEbfgs MATLAB/Single Molecule Fitting 2D/testdata.mat",
    "mlebfgs_as3d": "MLEbfgs MATLAB/Single Molecule Fitting AS3D/test_data_3D.mat",
    "mlewt_2d": "MLEwt MATLAB/Single Molecule Fitting 2D/test_data_2D.mat",
    "mlewt_as3d": "MLEwt MATLAB/Single Molecule Fitting AS3D/test_data_3D.mat",
    "multi_2e_2d": "MLEwt MATLAB/Two emitter fitting BFGS 2D/Multi_fit_testdata.mat",
    "multi_3e_2d": "MLEwt MATLAB/Three emitter fitting BFGS 2D/Multi_fit_testdata.mat",
    "multi_2e_3d": "MLEwt MATLAB/Two emitter fitting BFGS 3D/Multi_fit_testdata.mat",
    "tubulin_3d": "MLEwt MATLAB/Two emitter fitting BFGS 3D/3d multi fit test data/TestData3D.mat",
}


def main():
    if not os.path.isdir(REF):
        print("reference tree not present; fixtures already committed")
        return 0
    rois = {}
    for tag, rel in SOURCES.items():
        d = sio.loadmat(os.path.join(REF, rel))
        for k, v in sorted(d.items()):
            if k.startswith("__") or not isinstance(v, np.ndarray) or v.dtype != np.uint16:
                continue
            rois["%s__%s" % (tag, k)] = v
    np.savez_compressed(os.path.join(OUT, "mat_rois.npz"), **rois)
    d = sio.loadmat(os.path.join(REF, "MLEwt MATLAB/spline intorplation manual/testdata.mat"))
    np.savez_compressed(os.path.join(OUT, "spline_meandata.npz"), MeanData1=d["MeanData1"], MeanData2=d["MeanData2"])
    print("wrote", len(rois), "ROI arrays")
    return 0


if __name__ == "__main__":
    sys.exit(main())
