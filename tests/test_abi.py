"""The C-ABI library loads and exports every entry point include/qcstorm.h declares.

No compute calls: these run without a GPU.
"""
import ctypes as C
import os
import re

import qcstorm as Q

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "qcstorm.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^[A-Za-z_][\w \*]*?\b(qcs_\w+)\s*\(", text, flags=re.M)))


def test_header_parses():
    names = declared_functions()
    assert "qcs_fit_dev" in names and "qcs_session_feed" in names and len(names) >= 30


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(Q.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert missing == []


def test_binding_covers_header():
    assert sorted(Q.PROTOTYPES) == declared_functions()
    assert Q.missing_symbols() == []


def test_host_helpers_without_gpu():
    p = Q.default_params(ImageWidth=2048, ImageHigh=2048, PixelZoom=10.0)
    Q.lib().qcs_params_update_sr_size(C.byref(p))
    assert (p.SRImageWidth, p.SRImageHigh) == (20480, 20480)
    p = Q.default_params(ImageWidth=1021, ImageHigh=513, PixelZoom=5.0)
    Q.lib().qcs_params_update_sr_size(C.byref(p))
    assert (p.SRImageWidth, p.SRImageHigh) == (5108, 2568)   # multiples of 4 (bfgs_base.cu:89-98)
    assert Q.lib().qcs_version().startswith(b"qcstorm")
    assert Q.roi_words(7) == 56


def test_result_file_name_without_gpu():
    """SetLocDataFileName (IJ/LocResource.cpp:160-179): CString::TrimRight(".tif") drops every
    trailing '.', 't', 'i', 'f', then the suffix follows the mode."""
    p = Q.default_params()
    assert Q.result_file_name(p, "D:/data/cells.tif") == "D:/data/cells_result2D7_S.txt"
    assert Q.result_file_name(p, "stif.tif") == "s_result2D7_S.txt"          # trailing set characters go too
    assert Q.result_file_name(p, "movie") == "movie_result2D7_S.txt"
    q = Q.default_params(LocType=Q.LOCTYPE_AS3D, ROISize=11, MultiEmitterFitEn=1, ConsecFitEn=1,
                         ConsecFit_DistanceTh_nm=79.6)
    assert Q.result_file_name(q, "a.tif") == "a_result3D11_M_Consec80nm.txt"
    buf = C.create_string_buffer(8)
    n = Q.lib().qcs_result_file_name(C.byref(p), b"cells.tif", buf, 8)
    assert n == len("cells_result2D7_S.txt") and buf.value == b"cells_r"   # truncated, terminated
