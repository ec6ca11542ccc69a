"""The slump pass of the restatement (shed_pass, slump, commit_tmp;
film.wgsl:563-801, 866-876), stage by stage on hand-built fields, against
what the definitions fix:

- a straight front of a uniform film: the cells next to the dry side carry
  offset (1, 0), so their cone cap is tan(theta) dx, or the yield stress's
  steeper (tau / (rho h) - |g_xy|) / g_z times dx when that is larger; the
  dry cells within radius sqrt 5 sum to the +x direction, so each front cell
  sheds min(h - cap, h / 2) onto its dry neighbour alone (bilinear weights
  (1, 0, 0, 0)), which inherits its stratum; the cells behind the front
  shed nothing (their downhill neighbour is as high as they are);
- below twice the meniscus height nothing sheds, and the pass leaves every
  height and face volume bit-unchanged;
- on a spiky random canvas the total height is conserved, and every cell's
  height change is minus the divergence of the face volumes the pass adds to
  acc (each face's net transfer, once, on its owning cell)."""
import math

import numpy as np
import pytest

import oracle as orc
import scenes
from helpers import make_pair

L = orc.lib()


def film(w, h, tilt=0.0):
    s = scenes.blobs_settings(w, tilt_deg=tilt)
    s.slump = 1
    _, o = make_pair(w, h, s, None, gpu=False)
    o.update_params(1.0 / 60.0, 8)
    return o


def set_film(o, hgt, ty=0.0, mu=2.0):
    """Heights, one stratum each, rheology; tiles marked and listed."""
    hgt = np.asarray(hgt, np.float32)
    o.buf("h")[...] = hgt
    o.buf("strata_h")[...] = 0.0
    o.buf("strata_h")[..., 0] = hgt
    rh = o.buf("rheo")
    rh[..., 0] = np.where(hgt > 0, mu, 1.0)
    rh[..., 1] = np.where(hgt > 0, ty, 0.0)
    pig = o.buf("strata_pig")
    pig[...] = 0
    pig[..., 0, 0] = np.where(hgt > 0, 0x3C00, 0)   # pigment 0 = 1.0 (f16) in the wet cells' stratum
    o.buf("acc")[...] = 0.0
    o.buf("tile_moved")[...] = 1
    L.of_k_mark_compact(o._p)
    for _ in range(40):               # offsets to the nearest dry cell, one ring per pass
        L.of_k_nose(o._p)


def slump(o):
    L.of_k_shed(o._p)
    L.of_k_slump(o._p)
    L.of_k_commit_tmp(o._p)


@pytest.mark.parametrize("ty", [0.0, 30.0])
def test_straight_front_sheds_its_excess_onto_the_dry_side(ty):
    w = h = 32
    o = film(w, h)
    H = 3e-3
    hgt = np.zeros((h, w), np.float32)
    hgt[:, :16] = H
    set_film(o, hgt, ty=ty)
    P = o.params
    dx, tan_t, hmin = float(P.dx), float(P.ein[0]), 2.0 * float(P.ein[1])
    assert H > hmin
    gxy = math.hypot(float(P.g[0]), float(P.g[1]))
    need = max(ty / (float(P.phys[0]) * H) - gxy, 0.0) / max(float(P.g[2]), 1e-3)
    cap = max(tan_t, need) * dx
    amount = min(H - cap, 0.5 * H)
    assert amount > 0
    for j in range(2, h - 2):       # rows clear of the canvas edge (outside counts as dry)
        c = j * w + 15
        assert L.of_fn_shed_excess(o._p, c) == pytest.approx(H - cap, rel=1e-5)
    before_acc = o.buf("acc").copy()
    slump(o)
    hn = o.buf("h")
    rows = slice(2, h - 2)
    assert np.allclose(hn[rows, 15], H - amount, rtol=1e-5)
    assert np.allclose(hn[rows, 16], amount, rtol=1e-5)
    assert np.all(hn[rows, :15] == np.float32(H))          # behind the front: nothing moves
    assert np.all(hn[rows, 17:] == 0.0)
    # the new wet cells inherit the donor's stratum and rheology
    assert np.allclose(o.buf("strata_h")[rows, 16, 0], amount, rtol=1e-5)
    assert np.all(o.buf("strata_pig")[rows, 16, 0, 0] == 0x3C00)
    assert np.all(o.buf("rheo")[rows, 16, 1] == np.float32(ty))
    # the +x face of each front cell carries amount * dx
    dacc = o.buf("acc")[rows, 15, 0] - before_acc[rows, 15, 0]
    assert np.allclose(dacc, amount * dx, rtol=1e-5)


def test_nothing_sheds_below_twice_the_meniscus_height():
    w = h = 32
    o = film(w, h)
    hmin = 2.0 * float(o.params.ein[1])
    rng = np.random.default_rng(3)
    hgt = np.where(rng.random((h, w)) < 0.6, rng.uniform(0.05, 0.95, (h, w)) * hmin, 0.0).astype(np.float32)
    set_film(o, hgt, ty=5.0)
    h0, a0 = o.buf("h").copy(), o.buf("acc").copy()
    slump(o)
    assert np.array_equal(o.buf("h"), h0)
    assert np.array_equal(o.buf("acc"), a0)


def spiky(w, h, seed):
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:h, 0:w] + 0.5
    f = np.zeros((h, w))
    for _ in range(10):
        cx, cy, r = rng.uniform(0, w), rng.uniform(0, h), rng.uniform(0.8, 2.5)
        f += rng.uniform(2e-3, 9e-3) * np.exp(-((xx - cx) ** 2 + (yy - cy) ** 2) / (2 * r * r))
    return np.where(f > 2e-5, f, 0.0).astype(np.float32)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_mass_and_face_volumes_balance(seed):
    w = h = 48
    o = film(w, h, tilt=10.0)
    set_film(o, spiky(w, h, seed), ty=2.0)
    h0 = o.buf("h").astype(np.float64)
    a0 = o.buf("acc").astype(np.float64)
    slump(o)
    h1 = o.buf("h").astype(np.float64)
    da = (o.buf("acc").astype(np.float64) - a0) / float(o.params.dx)
    assert np.abs(h1 - h0).max() > 1e-4, "the canvas shed nothing"
    assert abs(h1.sum() - h0.sum()) <= 1e-6 * h0.sum()
    # height change = - (outflow through the owned faces +x, +y, +x+y, +x-y)
    #                 + (the same faces owned by the -x, -y, -x-y, -x+y neighbours)
    out = da.sum(-1)
    inn = np.zeros_like(out)
    inn[:, 1:] += da[:, :-1, 0]
    inn[1:, :] += da[:-1, :, 1]
    inn[1:, 1:] += da[:-1, :-1, 2]
    inn[:-1, 1:] += da[1:, :-1, 3]
    assert np.abs((h1 - h0) - (inn - out)).max() <= 1e-9
