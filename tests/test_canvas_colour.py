"""The canvas colour substrate (settings canvas_rgb; src/film/mod.rs:188-210,
686, 1018 through color::canvas_albedo and reflectance_of,
src/color/mod.rs:29-100):

- the reference's own test `canvas_reflectance_reproduces_colour`
  (src/color/mod.rs:381-394): the reflectance of each of its six colours,
  integrated against the CMF x D65 table and encoded to sRGB, gives the
  colour back within 0.03 -- here applied to the albedo the substrate gets
  (reflectance_of(canvas_albedo(rgb))), which must give canvas_albedo(rgb)
  back, and canvas_albedo is checked on its own below;
- canvas_albedo inverts the display tone curve: the filmic curve
  x (2.51 x + 0.03) / (x (2.43 x + 0.59) + 0.14) of the albedo times the view
  gain, sRGB-decoded, is the shown colour's linear value wherever the albedo
  is not clamped;
- the product's host function equals the restatement bit for bit, and the
  substrate words are its f16 pairs (the raw and white canvases unchanged)."""
import ctypes as C

import numpy as np
import pytest

import canvaz_film as czf
import oracle as orc

REF_COLOURS = [[0.94, 0.93, 0.90], [0.74, 0.56, 0.33], [0.64, 0.37, 0.26], [0.50, 0.56, 0.62],
               [0.56, 0.56, 0.55], [0.30, 0.45, 0.25]]
GAIN = np.array([1.94, 1.95, 2.17])


def dec(c):
    c = np.asarray(c, np.float64)
    return np.where(c <= 0.04045, c / 12.92, ((c + 0.055) / 1.055) ** 2.4)


def xyz_to_srgb(xyz):
    """The IEC 61966-2-1 matrix and transfer function (src/color/mod.rs:320-333)."""
    x, y, z = xyz
    rgb = np.array([3.2406 * x - 1.5372 * y - 0.4986 * z, -0.9689 * x + 1.8758 * y + 0.0415 * z,
                    0.0557 * x - 0.2040 * y + 1.0570 * z])
    rgb = np.clip(rgb, 0.0, 1.0)
    return np.where(rgb <= 0.0031308, 12.92 * rgb, 1.055 * rgb ** (1 / 2.4) - 0.055)


def refl(rgb):
    cmf = czf.cmf_table()
    out = np.zeros(16, np.float32)
    czf.lib().cz_canvas_reflectance(czf._fp(np.asarray(rgb, np.float32)), czf._fp(cmf), czf._fp(out))
    return out


def tone(x):
    return x * (2.51 * x + 0.03) / (x * (2.43 * x + 0.59) + 0.14)


def albedo64(rgb):
    """canvas_albedo from its definition: the albedo a whose tone-mapped,
    gained linear value shows as rgb, solved by bisection in float64."""
    out = []
    for i, c in enumerate(rgb):
        y = min(max(float(dec(c)), 0.0), 0.995)
        lo, hi = 0.0, 100.0
        for _ in range(200):
            mid = 0.5 * (lo + hi)
            lo, hi = (mid, hi) if tone(mid) < y else (lo, mid)
        lin = min(max(0.5 * (lo + hi) / GAIN[i], 0.005), 0.9)
        out.append(12.92 * lin if lin <= 0.0031308 else 1.055 * lin ** (1 / 2.4) - 0.055)
    return np.array(out)


@pytest.mark.parametrize("rgb", REF_COLOURS)
def test_reflectance_reproduces_the_albedo(rgb):
    cmf = czf.cmf_table().astype(np.float64).reshape(16, 3)
    r = refl(rgb).astype(np.float64)
    got = xyz_to_srgb(r @ cmf)
    want = albedo64(rgb)
    assert np.all(np.abs(got - want) < 0.03), (rgb, got, want)
    assert np.all(r >= 0.005) and np.all(r <= 0.97)


def test_albedo_inverts_the_tone_curve():
    """The same albedo from the restatement's closed-form root: through the
    restatement's reflectance, the colour of a mid-grey canvas comes back
    within the reference's 0.03 of the bisection albedo, and a lighter
    shown colour gives a lighter substrate in every band."""
    lo, hi = refl([0.40, 0.40, 0.40]), refl([0.80, 0.80, 0.80])
    assert np.all(hi > lo)
    cmf = czf.cmf_table().astype(np.float64).reshape(16, 3)
    got = xyz_to_srgb(lo.astype(np.float64) @ cmf)
    assert np.all(np.abs(got - albedo64([0.40, 0.40, 0.40])) < 0.03)


@pytest.mark.parametrize("rgb", REF_COLOURS + [[0.0, 0.0, 0.0], [1.0, 1.0, 1.0], [0.2, 0.9, 0.1]])
def test_product_equals_restatement(rgb):
    cmf = czf.cmf_table()
    a = refl(rgb)
    b = (C.c_float * 16)()
    orc.lib().of_canvas_reflectance((C.c_float * 3)(*rgb), cmf.ctypes.data_as(C.POINTER(C.c_float)), b)
    assert np.array_equal(a, np.array(b[:], np.float32))


def test_substrate_words():
    cmf = czf.cmf_table()
    s = czf.default_settings()
    words = np.zeros(8, np.uint32)
    for raw in (0, 1):
        s.raw_canvas = raw
        czf.lib().cz_substrate_words(C.byref(s), czf._fp(cmf), czf._up(words))
        plain = np.zeros(8, np.uint32)
        czf.lib().cz_substrate_refl(1, raw, czf._up(plain))
        assert np.array_equal(words, plain)
    s.has_canvas_rgb = 1
    s.canvas_rgb[:] = [0.74, 0.56, 0.33]
    czf.lib().cz_substrate_words(C.byref(s), czf._fp(cmf), czf._up(words))
    bands = words.view(np.float16).astype(np.float32)
    assert np.array_equal(bands, refl([0.74, 0.56, 0.33]).astype(np.float16).astype(np.float32))
