"""Pins of the restated film physics against what the mathematics fixes:
closed-form film fluxes, the band-flux telescoping identity, Kubelka-Munk
limits, the 9-point Laplacian on quadratics, the plug mean of a uniform film,
the cure-to-rheology closed form and the tile wake rule."""
import ctypes as C
import math

import numpy as np
import pytest

import canvaz_film as czf
import oracle as orc
import scenes
from helpers import make_pair

L = orc.lib()


def hb(n, g, h, mu, ty):
    y = C.c_float()
    up = C.c_float()
    q = L.of_hb_flux(n, g, h, mu, ty, C.byref(y), C.byref(up))
    return q, y.value, up.value


@pytest.mark.parametrize("h", [1e-5, 2e-4, 1e-3, 3e-3])
def test_newtonian_film_flux(h):
    """tau = 0, n = 1: q = g h^3 / (3 mu) (Nusselt film)."""
    g, mu = 2539.0, 3.7
    q, y, up = hb(1.0, g, h, mu, 0.0)
    assert q == pytest.approx(g * h ** 3 / (3 * mu), rel=3e-6)
    assert y == pytest.approx(h, rel=1e-7)


@pytest.mark.parametrize("n", [0.4, 0.55, 0.8, 0.95])
def test_power_law_film_flux(n):
    """tau = 0: q = n / (2n + 1) (g / mu)^(1/n) h^((2n+1)/n)."""
    g, mu, h = 1800.0, 12.0, 1.2e-3
    q, _, _ = hb(n, g, h, mu, 0.0)
    ref = n / (2 * n + 1) * (g / mu) ** (1 / n) * h ** ((2 * n + 1) / n)
    assert q == pytest.approx(ref, rel=2e-5)


@pytest.mark.parametrize("ty", [0.1, 1.0, 2.0])
def test_bingham_plug_flux(ty):
    """n = 1, tau > 0: y = h - tau / g, q = g / (2 mu) y^2 (h - y / 3); no flow below the yield depth."""
    g, mu, h = 2500.0, 4.0, 1.5e-3
    q, y, _ = hb(1.0, g, h, mu, ty)
    yy = h - ty / g
    assert y == pytest.approx(yy, rel=1e-5)
    assert q == pytest.approx(g / (2 * mu) * yy * yy * (h - yy / 3), rel=2e-5)
    q0, y0, _ = hb(1.0, g, h, mu, 1.01 * g * h)
    assert q0 == 0.0 and y0 == 0.0


@pytest.mark.parametrize("n", [1.0, 0.6])
def test_band_flux_telescopes(n):
    """The band fluxes of any partition of [0, h] sum to the film flux."""
    rng = np.random.default_rng(7)
    for _ in range(20):
        g, mu, h = rng.uniform(500, 5000), rng.uniform(0.5, 30), rng.uniform(1e-4, 3e-3)
        ty = rng.uniform(0, 0.5 * g * h)
        q, y, up = hb(n, g, h, mu, ty)
        cuts = np.sort(np.concatenate([[0.0], rng.uniform(0, h, 3), [h]])).astype(np.float32)
        s = sum(L.of_band_flux(n, up, y, float(a), float(b)) for a, b in zip(cuts[:-1], cuts[1:]))
        assert s == pytest.approx(q, rel=2e-4, abs=1e-12)


def test_km_layer_limits():
    R, T = C.c_float(), C.c_float()
    for k, s in [(0.3, 2.0), (5.0, 1.0), (0.01, 10.0)]:
        L.of_km_layer(k, s, 1e3, C.byref(R), C.byref(T))
        a = k / s
        assert R.value == pytest.approx(1 + a - math.sqrt(a * a + 2 * a), rel=2e-6)
        assert T.value < 1e-6
        L.of_km_layer(k, s, 0.0, C.byref(R), C.byref(T))
        assert (R.value, T.value) == (0.0, 1.0)
    # composite over a black substrate is the layer's own reflectance
    assert L.of_composite(0.3, 0.5, 0.0) == pytest.approx(0.3)
    # a clear layer (R = 0, T = 1) leaves the substrate unchanged
    assert L.of_composite(0.0, 1.0, 0.42) == pytest.approx(0.42)


def tiny_film(w=32, h=32, tilt=0.0):
    s = scenes.blobs_settings(w, tilt_deg=tilt)
    _, o = make_pair(w, h, s, gpu=False)
    return o


def activate_all(o):
    o.buf("tile_moved")[...] = 1
    o.buf("h_frame")[...] = -1.0
    L.of_k_mark_compact(o._p)
    assert o.active_count() > 0


def test_laplacian_exact_on_quadratic():
    """pres.x = -sigma lap(s) + Pi(h); the 2/3, 1/6 stencil is exact on x^2 + y^2."""
    o = tiny_film()
    dx = o.params.dx
    W = H = 32
    A = 3.0
    j, i = np.mgrid[0:H, 0:W].astype(np.float64)
    surf = A * (((i - 16) * dx) ** 2 + ((j - 16) * dx) ** 2) + 2e-3
    o.buf("m")[...] = 0.0
    o.buf("h")[...] = surf.astype(np.float32)
    o.buf("delta")[...] = 0
    activate_all(o)
    L.of_k_laplacian(o._p)
    p = o.params
    pres = o.buf("pres")[2:-2, 2:-2, 0].astype(np.float64)
    hh = o.buf("h")[2:-2, 2:-2].astype(np.float64)
    disj = p.ein[2] * np.maximum(1 - hh / p.cure[3], 0) ** 2
    lap = (disj - pres) / p.phys[1]
    assert np.allclose(lap, 4 * A, rtol=2e-3)


def test_plug_mean_of_uniform_film():
    """On a uniform yield-stress film the 7x7 mean is the cell's own hb_flux / h."""
    o = tiny_film(tilt=30.0)
    hgt, mu, ty = 2e-3, 3.0, 4.0
    o.buf("h")[...] = hgt
    o.buf("rheo")[..., 0] = mu
    o.buf("rheo")[..., 1] = ty
    activate_all(o)
    L.of_k_plug(o._p)
    p = o.params
    gb = np.float32(p.phys[0]) * np.float32(math.hypot(p.g[0], p.g[1]))
    q, _, _ = hb(p.phys[2], float(gb), hgt, mu, ty)
    assert np.allclose(o.buf("tmp"), q / hgt, rtol=1e-6)


def test_cure_rheology_closed_form():
    """One stratum: mu = exp(ln mu0 + 3 ln(1 - thin) + lnmu_cure cb^2),
    tau = tau0 (1 - thin)^4 + tau_cure cb^2 (shaders/cure.wgsl:137-192)."""
    o = tiny_film(8, 8)
    rng = np.random.default_rng(11)
    frac = np.zeros((8, 8, 16), np.float32)
    frac[..., :12] = rng.dirichlet(np.full(12, 0.5), (8, 8))
    thin = rng.uniform(0, 0.5, (8, 8)).astype(np.float32)
    tau0 = rng.uniform(0.1, 40, (8, 8)).astype(np.float32)
    lnmu0 = rng.uniform(-1, 3, (8, 8)).astype(np.float32)
    cure = rng.uniform(0, 0.9, (8, 8)).astype(np.float32)
    frac[..., 12], frac[..., 13], frac[..., 14] = thin, tau0, lnmu0 + 10
    o.buf("h")[...] = 1e-3
    o.buf("strata_h")[..., 0] = 1e-3
    o.buf("strata_pig")[:, :, 0, :] = scenes.pack_pig(frac)
    o.buf("strata_cure")[..., 0] = cure
    L.of_k_cure(o._p)
    p = o.params
    q = frac.astype(np.float16).astype(np.float64)
    cb = np.minimum(cure.astype(np.float64) + p.cure[0] * p.frame_dt, 1.0)
    tf = np.clip(q[..., 12], 0, 0.95)
    mu = np.exp((q[..., 14] - 10) + 3 * np.log(1 - tf) + p.cure[2] * cb * cb)
    tau = q[..., 13] * (1 - tf) ** 4 + p.cure[1] * cb * cb
    r = o.buf("rheo").astype(np.float64)
    assert np.allclose(r[..., 0], mu, rtol=2e-5)
    assert np.allclose(r[..., 1], tau, rtol=2e-5)


def test_tile_wake_rule():
    """compact_tiles (film.wgsl:1393-1425): a tile is active if it or a
    neighbour whose wet band faces it moved; written here as shifts of the
    flag raster."""
    o = tiny_film(96, 80)
    rng = np.random.default_rng(3)
    tw, th = 6, 5
    for trial in range(10):
        hgt = np.zeros((80, 96), np.float32)
        mask = rng.random((80, 96)) < 0.02
        hgt[mask] = 1e-4
        o.buf("h")[...] = hgt
        o.buf("h_frame")[...] = np.where(rng.random((80, 96)) < 0.5, hgt, 0.0)
        o.buf("tile_moved")[...] = (rng.random(tw * th) < 0.1).astype(np.uint32)
        tm = o.buf("tile_moved").copy().reshape(th, tw)
        L.of_k_mark_compact(o._p)
        fl = o.buf("tile_flags").reshape(th, tw)
        # flags from the definitions (film.wgsl:1365-1391)
        exp_bits = np.zeros((th, tw), np.uint32)
        hf = o.buf("h_frame")
        for ty in range(th):
            for tx in range(tw):
                blk = hgt[ty * 16:(ty + 1) * 16, tx * 16:(tx + 1) * 16]
                fr = hf[ty * 16:(ty + 1) * 16, tx * 16:(tx + 1) * 16]
                wet = blk > 0
                b = 0
                if wet.any():
                    b |= 1
                    b |= 2 if wet[:, :8].any() else 0
                    b |= 4 if wet[:, 8:].any() else 0
                    b |= 8 if wet[:8, :].any() else 0
                    b |= 16 if wet[8:, :].any() else 0
                    b |= 64 if (np.abs(blk - fr)[wet] > 2e-6).any() else 0
                if tm[ty, tx]:
                    b |= 64
                exp_bits[ty, tx] = b
        assert np.array_equal(fl & 95, exp_bits)
        moved = (exp_bits & 64) != 0
        act = moved.copy()
        pad = np.zeros((th + 2, tw + 2), np.uint32)
        pad[1:-1, 1:-1] = exp_bits
        for dy in (-1, 0, 1):
            for dxi in (-1, 0, 1):
                if dxi == 0 and dy == 0:
                    continue
                nb = pad[1 + dy:1 + dy + th, 1 + dxi:1 + dxi + tw]
                need = np.ones_like(nb, bool)
                if dxi == 1:
                    need &= (nb & 2) != 0
                if dxi == -1:
                    need &= (nb & 4) != 0
                if dy == 1:
                    need &= (nb & 8) != 0
                if dy == -1:
                    need &= (nb & 16) != 0
                act |= need & ((nb & 64) != 0)
        assert np.array_equal((fl & 32) != 0, act)
        assert np.array_equal(o.buf("tile_list")[:o.active_count()], np.flatnonzero(act.reshape(-1)))
