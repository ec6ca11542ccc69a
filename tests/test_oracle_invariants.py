"""Invariants of the restated frame on seeded scenes: mass bookkeeping of the
mass report (src/main.rs:1725-1726,1815), non-negative heights, strata
summing to the height, pigment fractions staying convex combinations, and
determinism (the restatement pins the reference's racy dispatches)."""
import numpy as np
import pytest

import scenes
from helpers import make_pair, oracle_capsule


def run_blobs(frames, seed=4, w=64, tilt=20.0, deposits=True):
    s = scenes.blobs_settings(w, tilt_deg=tilt)
    st = scenes.blobs_state(w, w, seed=seed, nblobs=8, sigma_cells=(2.0, 7.0), tau=(0.05, 2.0), mu=(0.5, 5.0))
    _, o = make_pair(w, w, s, st, gpu=False)
    dx = o.params.dx
    h0 = float(o.buf("h").astype(np.float64).sum()) * dx * dx
    pig = scenes.jet_paint()
    for f in range(frames):
        deps = []
        if deposits and f < frames // 2:
            c = (0.3 * w * dx + 0.4 * w * dx * f / frames, 0.5 * w * dx)
            deps = [oracle_capsule(c, (c[0] + 2 * dx, c[1]), 3 * dx, 2e-9, pig, dx)]
        o.step(deps)
    return o, h0, dx


def test_mass_bookkeeping():
    o, h0, dx = run_blobs(24)
    r = o.mass_report()
    assert r.nan_cells == 0
    wet = r.wet_ml * 1e-6
    cured = r.cured_ml * 1e-6
    dep = r.deposited_ml * 1e-6
    assert dep > 0
    assert abs(wet + cured - h0 - dep) <= 2e-6 * (h0 + dep)


def test_heights_and_strata():
    o, _, _ = run_blobs(16)
    h = o.buf("h")
    assert (h >= 0).all() and np.isfinite(h).all()
    sh = o.buf("strata_h").astype(np.float64)
    wet = h > 0
    ssum = sh.sum(-1)
    # every wet cell of an active tile carries strata whose sum is the height
    act = np.zeros_like(wet)
    tl = o.buf("tile_list")[:o.active_count()]
    tw = (o.w + 15) // 16
    for t in tl:
        ty, tx = divmod(int(t), tw)
        act[ty * 16:(ty + 1) * 16, tx * 16:(tx + 1) * 16] = True
    sel = wet & act & (ssum > 0)
    assert sel.any()
    assert np.allclose(ssum[sel], h[sel], rtol=1e-4)
    assert (sh >= 0).all()
    pig = o.buf("strata_pig").view(np.float16).astype(np.float32).reshape(o.h, o.w, 4, 16)
    conc = pig[..., :12]
    assert (conc >= 0).all()
    used = sh > 0
    # f16 storage: each of the 12 fractions carries a 2^-11 relative rounding per store
    assert (conc.sum(-1)[used] <= 1.0 + 1e-2).all()


def test_deterministic():
    a, _, _ = run_blobs(8, seed=9)
    b, _, _ = run_blobs(8, seed=9)
    for k in ("h", "strata_h", "strata_pig", "delta", "rheo", "tile_flags"):
        assert np.array_equal(a.buf(k), b.buf(k)), k


def test_uniform_film_at_rest_stays():
    """A film covering the whole flat, untilted canvas has no contact line
    and no driving gradient: it stays uniform to the bit."""
    w = 48
    s = scenes.blobs_settings(w, tilt_deg=0.0)
    st = scenes.blobs_state(w, w, seed=1, nblobs=0)
    st["m"][...] = 0.0
    st["h"][...] = 1e-3
    st["strata_h"][..., 0] = 1e-3
    frac = np.zeros(16, np.float32)
    frac[0] = 1.0
    frac[14] = np.log(5.0) + 10
    st["strata_pig"][:, :, 0, :] = scenes.pack_pig(frac)
    _, o = make_pair(w, w, s, st, gpu=False)
    for _ in range(3):
        o.step()
    h = o.buf("h")
    assert (h == np.float32(1e-3)).all()
