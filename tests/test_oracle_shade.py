"""The restated shade pass (shade.wgsl:29-83, km.wgsl:35-62) against what the
formulas fix: a bare canvas cell is the substrate under the Saunderson
surface correction mapped by CIE 1931 x D65 to linear sRGB; an optically
thick layer shows its KM infinite-thickness reflectance R_inf = 1 + a -
sqrt(a^2 + 2a), a = K/S, whatever lies beneath; roughness and the detail
channels follow their definitions; the debug view colours the stack cases."""
import numpy as np
import pytest

import canvaz_film as czf
import oracle as orc

XYZ_TO_RGB = np.array([[3.2406, -1.5372, -0.4986], [-0.9689, 1.8758, 0.0415], [0.0557, -0.2040, 1.0570]])


def f16(bits):
    return np.asarray(bits, np.uint16).view(np.float16).astype(np.float64)


def film(w=4, h=1, debug=0):
    s = czf.default_settings()
    s.width_m, s.height_m = w * 1e-3, h * 1e-3
    s.debug_view = debug
    o = orc.OracleFilm(w, h, s, czf.default_pigment_table(), czf.cmf_table())
    o.update_params(1.0 / 60.0, 4)
    for k in ("h", "m", "strata_h", "strata_cure", "strata_pig", "lay"):
        o.buf(k)[...] = 0
    return o


def cells(o, name):
    """Per-cell view of an oracle buffer (row-major cells)."""
    b = o.buf(name)
    return b.reshape((o.w * o.h,) + b.shape[2:])


def set_refl(o, c, rho):
    half = np.float16(rho).view(np.uint16).astype(np.uint32)
    cells(o, "sub_refl")[c] = half | (half << 16)


def expected_rgb(r16, wet):
    k2 = 0.4 * (1 - wet) + 0.6 * wet
    rs = 0.96 * (1 - k2) * r16 / np.maximum(1 - k2 * r16, 1e-3)
    cmf = czf.cmf_table().astype(np.float64).reshape(16, 3)
    return np.maximum(XYZ_TO_RGB @ (rs @ cmf), 0.0)


def test_bare_canvas_is_the_corrected_substrate():
    o = film()
    rhos = [0.0, 0.25, 0.5, 0.82]
    for c, rho in enumerate(rhos):
        set_refl(o, c, rho)
    cells(o, "lay")[2] = 3 | (32768 << 8)
    cells(o, "m")[3] = 80e-6
    color, detail = o.shade()
    for c, rho in enumerate(rhos):
        want = expected_rgb(np.full(16, float(np.float16(rho))), 0.0)
        np.testing.assert_allclose(f16(color[0, c, :3]), want, rtol=2e-3, atol=1e-4)
    painted = lambda m: np.clip((m - 60e-6) / 90e-6, 0, 1) ** 2 * (3 - 2 * np.clip((m - 60e-6) / 90e-6, 0, 1))
    assert f16(color[0, 0, 3]) == pytest.approx(0.85, abs=1e-3)            # bare canvas roughness
    p = painted(80e-6)
    assert f16(color[0, 3, 3]) == pytest.approx(0.85 * (1 - p) + 0.35 * p, abs=1e-3)
    assert f16(detail[0, 2, 0]) == pytest.approx(3 + 0.5, abs=1e-3)         # method + turn
    assert f16(detail[0, 3, 3]) == pytest.approx(20e-6, abs=1e-7)           # thickness (an f16 subnormal)
    assert f16(detail[0, 3, 1]) == pytest.approx(p, abs=1e-3)               # cure_top of a dry painted cell


def test_thick_layer_shows_r_infinity():
    o = film(w=3)
    pt = czf.default_pigment_table().astype(np.float64).reshape(12, -1)
    for c, rho in enumerate((0.05, 0.5, 0.9)):   # any substrate
        set_refl(o, c, rho)
        cells(o, "h")[c] = 5e-3
        cells(o, "strata_h")[c] = (5e-3, 0, 0, 0)
        cells(o, "strata_cure")[c] = (0.25, 0, 0, 0)
        frac = np.zeros(16, np.float32)
        frac[7] = 1.0                           # one pigment
        bits = frac.astype(np.float16).view(np.uint16).astype(np.uint32)
        cells(o, "strata_pig")[c, 0] = bits[0::2] | (bits[1::2] << 16)
    color, detail = o.shade()
    k, s = pt[7, :16], pt[7, 16:32]
    a = k / s
    r_inf = 1 + a - np.sqrt(a * a + 2 * a)
    want = expected_rgb(r_inf, 0.75)
    for c in range(3):
        np.testing.assert_allclose(f16(color[0, c, :3]), want, rtol=3e-3, atol=1e-4)
        assert f16(detail[0, c, 1]) == pytest.approx(0.25, abs=1e-3)
        assert f16(color[0, c, 3]) == pytest.approx(0.06 + (0.35 - 0.06) * 0.5, abs=1e-3)  # sqrt(0.25)


def test_debug_view_cases():
    o = film(w=4, debug=1)
    one = np.zeros(16, np.float32)
    one[2] = 1.0
    bits = one.astype(np.float16).view(np.uint16).astype(np.uint32)
    packed = bits[0::2] | (bits[1::2] << 16)
    cells(o, "h")[:] = 1e-4
    cells(o, "strata_h")[0] = 0                                   # empty stack: blue
    cells(o, "strata_h")[1] = (1e-4, 0, 0, 0)                     # no pigment: red
    cells(o, "strata_h")[2] = (5e-5, 0, 0, 0)                     # sum off h: green
    cells(o, "strata_pig")[2, 0] = packed
    cells(o, "strata_h")[3] = (5e-5, 5e-5, 0, 0)                  # two strata: 0.3 + 0.15 * 2
    cells(o, "strata_pig")[3, 0] = packed
    cells(o, "strata_pig")[3, 1] = packed
    color, _ = o.shade()
    assert list(f16(color[0, 0, :3])) == [0, 0, 1]
    assert list(f16(color[0, 1, :3])) == [1, 0, 0]
    assert list(f16(color[0, 2, :3])) == [0, 1, 0]
    np.testing.assert_allclose(f16(color[0, 3, :3]), [0.6] * 3, atol=1e-3)
