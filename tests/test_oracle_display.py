"""Pinning the oracle's colour-coded 3D image (oracle/orc_display.c): the HSV colours at the hue
anchors, the colour-mode flip, the value cap and black background against their definitions, and
the 97% brightness level against a histogram with a known cumulative distribution."""
import numpy as np
import pytest

import oracle_lib as O
import qcstorm as Q


def _px(params, intensity, z, scale, mode=Q.RGB_RGB):
    img = np.array([[intensity]], np.float32)
    zimg = np.array([[intensity * z]], np.float32)
    return tuple(int(v) for v in O.display_rgb(params, img, zimg, scale, mode)[0, 0])


def test_hue_anchors_and_modes():
    p = Q.default_params(MinZDepth=-500.0, MaxZDepth=500.0, ColorMode_3D=0)   # blue (min z) to red (max z)
    assert _px(p, 100.0, -500.0, 100.0) == (0, 0, 255)     # hue 240: blue at full value (100 * 2.55 = 255)
    assert _px(p, 100.0, 0.0, 100.0) == (0, 255, 0)        # hue 120: green
    assert _px(p, 100.0, 500.0, 100.0) == (255, 0, 0)      # hue 0: red
    assert _px(p, 100.0, 900.0, 100.0) == (255, 0, 0)      # depth clamped to the range
    q = Q.default_params(MinZDepth=-500.0, MaxZDepth=500.0, ColorMode_3D=1)
    assert _px(q, 100.0, -500.0, 100.0) == (255, 0, 0)
    # hue 60 (z = 250 in mode 0): yellow
    assert _px(p, 100.0, 250.0, 100.0) == (255, 255, 0)
    # value = intensity * 120 / scale, capped at 100; 50 -> 127.5 + 0.5 -> 128
    assert _px(p, 50.0, 500.0, 120.0) == (128, 0, 0)
    assert _px(p, 1e6, 500.0, 120.0) == (255, 0, 0)
    # encodings: BGR swaps, 4-byte forms add alpha 0xff; zero intensity is black
    assert _px(p, 100.0, 500.0, 100.0, Q.RGB_BGR) == (0, 0, 255)
    assert _px(p, 100.0, 500.0, 100.0, Q.RGB_BGRA) == (0, 0, 255, 255)
    assert _px(p, 0.0, 500.0, 100.0, Q.RGB_RGBA) == (0, 0, 0, 255)


def test_brightness_level():
    h = np.zeros(2048, np.int32)
    h[:100] = 1                     # 100 pixels: 97% is reached after bin 96
    lvl, _ = O.display_max_intensity(hist=h)
    assert lvl == pytest.approx(97 * 12000.0 / 2048, rel=1e-6)
    # from an image: values v land in bin trunc(v / gap - 1); values in (0, 2 gap) in bin 0
    gap = 12000.0 / 2048
    img = np.array([0.0, 0.5 * gap, 1.5 * gap, 2.5 * gap, 20000.0], np.float32)
    _, hh = O.display_max_intensity(img)
    assert hh[0] == 2 and hh[1] == 1 and hh.sum() == 3
