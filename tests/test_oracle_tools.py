"""The restated tool passes (brush_scrape, shaders/brush.wgsl:33-122; knife,
shaders/knife.wgsl:56-128; the contact binning of src/film/mod.rs:1044-1084)
pinned by what their rules fix on a flat single-stratum film: the caps each
contact kind leaves, the cells a contact or the blade reaches (brute force
over segment distance and the blade's parallelogram, computed here in
float64), the laying direction as a closed-form angle, the reservoir and
the scraps accounting for every cubic metre the film lost, the scrap order,
and the coverage fraction of the knife's fill against ratio^0.6 (which pins
the cell hash's uniformity and the exponent)."""
import math

import numpy as np
import pytest

import cz_types
import scenes
from helpers import make_pair

W, H = 96, 80
DX = 1.2 / 2048
PARTS = ("h", "m", "rheo", "strata_h", "strata_pig", "strata_cure", "sub_refl", "lay")


def flat_state(h0, pig_index=0, yield_stress=2.0, ln_mu=10.5, dry_cols=0):
    frac = np.zeros(16, np.float32)
    frac[pig_index] = 1.0
    frac[13] = yield_stress
    frac[14] = ln_mu
    hgt = np.full((H, W), h0, np.float32)
    if dry_cols:
        hgt[:, :dry_cols] = 0.0
    pig = np.zeros((H, W, 4, 8), np.uint32)
    pig[:, :, 0, :] = scenes.pack_pig(frac)
    pig[hgt == 0] = 0
    sh = np.zeros((H, W, 4), np.float32)
    sh[:, :, 0] = hgt
    rheo = np.zeros((H, W, 2), np.float32)
    rheo[:, :, 0] = 1.0
    sub = np.zeros((H, W, 8), np.uint32)
    return {"h": hgt, "m": np.zeros((H, W), np.float32), "rheo": rheo, "strata_h": sh, "strata_pig": pig,
            "strata_cure": np.zeros((H, W, 4), np.float32), "sub_refl": sub, "lay": np.zeros((H, W), np.uint32)}


def film(state):
    s = scenes.blobs_settings(W)
    s.height_m = H * DX
    _, o = make_pair(W, H, s, state, gpu=False)
    return o


def contact(p0, p1, r, zb, bristle=0, kind=0):
    c = cz_types.Contact()
    c.p0[:] = p0
    c.p1[:] = p1
    c.r, c.zb, c.bristle, c.pad = r, zb, bristle, kind
    return c


def centres():
    ys, xs = np.mgrid[0:H, 0:W]
    return (xs + 0.5) * DX, (ys + 0.5) * DX


def seg_dist(px, py, p0, p1):
    ab = np.array(p1, np.float64) - p0
    l2 = ab @ ab
    t = np.clip(((px - p0[0]) * ab[0] + (py - p0[1]) * ab[1]) / l2, 0, 1) if l2 > 0 else 0.0
    return np.hypot(px - (p0[0] + ab[0] * t), py - (p0[1] + ab[1] * t))


def turn16(dx_, dy_):
    return int(np.float32((math.atan2(dy_, dx_) / (2 * math.pi) + 0.5)) * np.float32(65535.0))


@pytest.mark.parametrize("kind,zb,cap", [
    (0, 4e-5, 8e-5),    # bristle: cap = max(zb - m, LEAVE_FILM), a quarter of the excess taken
    (0, 1.5e-4, 1.5e-4),
    (1, 3e-5, 8e-5),    # roller: cap = max(zb, LEAVE_FILM), all the excess
    (2, 3e-5, 3e-5),    # sponge: cap = max(zb, SPONGE_FILM)
])
def test_brush_caps_reach_and_lay(kind, zb, cap):
    h0 = np.float32(2e-4)
    o = film(flat_state(h0))
    before = {k: o.buf(k).copy() for k in PARTS}
    p0, p1, r = (30 * DX, 40 * DX), (60 * DX, 40 * DX), 6 * DX
    o.upload_contacts([contact(p0, p1, r, zb, bristle=7, kind=kind)])
    o.brush_pass()
    px, py = centres()
    inside = seg_dist(px, py, p0, p1) < r
    hn = o.buf("h")
    excess = np.float32(h0 - np.float32(cap))
    take = excess * np.float32(0.25) if kind == 0 else excess
    want = np.float32(h0 - take) if h0 > cap else h0
    assert np.all(hn[inside] == want)
    assert np.array_equal(hn[~inside], before["h"][~inside])
    assert np.array_equal(o.buf("strata_h")[inside][:, 0], hn[inside])
    method = 6 if kind in (1, 2) else 4
    lay = o.buf("lay")[inside]
    assert np.all(lay & 0xFF == method)
    d = (0.0, -30 * DX) if kind == 1 else (30 * DX, 0.0)  # roller: the stroke turned a quarter
    assert np.all(lay >> 8 == turn16(d[0] + 1e-9, d[1]))
    assert np.array_equal(o.buf("lay")[~inside], before["lay"][~inside])
    res = o.take_reservoir()
    lost = float(np.sum((before["h"] - hn).astype(np.float64))) * DX * DX * 1e15
    ncell = int(inside.sum())
    assert abs(int(res[7, 0]) - lost) <= ncell, (res[7, 0], lost)
    assert np.array_equal(res[7, 1], res[7, 0])      # one pigment, fraction 1
    assert np.all(res[7, 2:13] == 0)
    assert abs(int(res[7, 14]) - 0.01 * 2.0 * lost) <= ncell  # yield stress slot, scaled by 0.01
    assert np.all(np.delete(res, 7, axis=0) == 0)
    assert not np.any(o.take_reservoir())            # taken: the sums restart


def test_brush_skips_thin_cells_and_keeps_lowest_cap():
    h0 = np.float32(2e-4)
    o = film(flat_state(h0, dry_cols=40))
    lay0 = o.buf("lay").copy()
    p0, p1, r = (20 * DX, 40 * DX), (70 * DX, 40 * DX), 5 * DX
    o.upload_contacts([contact(p0, p1, r, 1.2e-4, bristle=3), contact(p0, p1, 2 * r, 1.0e-4, bristle=600)])
    o.brush_pass()
    px, py = centres()
    d = seg_dist(px, py, p0, p1)
    dry = np.zeros((H, W), bool)
    dry[:, :40] = True
    # dry cells (h <= SPONGE_FILM) are left alone, lay included
    assert np.all(o.buf("h")[dry] == 0) and np.array_equal(o.buf("lay")[dry], lay0[dry])
    both = (d < r) & ~dry
    want = np.float32(h0 - np.float32(np.float32(h0 - np.float32(1.0e-4)) * np.float32(0.25)))
    assert np.all(o.buf("h")[both] == want)  # the lower cap (bristle 600) wins
    # bristle 600 has no reservoir slot: paint taken, nothing recorded
    assert not np.any(o.take_reservoir())


def test_contacts_outside_or_not_finite_do_nothing():
    o = film(flat_state(np.float32(2e-4)))
    h0 = o.buf("h").copy()
    far = W * DX + 1.0
    o.upload_contacts([contact((far, far), (far + 0.01, far), 0.01, 0.0),
                       contact((float("nan"), 0.0), (0.0, 0.0), 0.01, 0.0),
                       contact((-1.0, -1.0), (-0.5, -1.0), 0.1, 0.0)])
    o.brush_pass()
    assert np.array_equal(o.buf("h"), h0)


def knife(a, b, dirv, sweep, height, fill=0.0, smear=0.0, bands=8, frame=5, pig_index=3):
    k = cz_types.Knife()
    k.a[:] = a
    k.b[:] = b
    k.dir[:] = dirv
    k.sweep, k.height, k.half_width = sweep, height, 0.01
    k.active, k.frame, k.bands = 1, frame, bands
    frac = np.zeros(16, np.float32)
    frac[pig_index] = 1.0
    for bi in range(8):
        k.pig[bi][:] = [int(v) for v in scenes.pack_pig(frac)]
    k.fill[:] = [fill] * 8
    k.smear[:] = [smear] * 8
    return k


def blade_cells(a, b, dirv, sweep):
    """(t, behind, u) of every cell centre, in float64: p = a + t (b - a) - behind dir."""
    px, py = centres()
    ab = np.array(b, np.float64) - a
    mv = np.array(dirv, np.float64) * sweep
    det = ab[0] * mv[1] - ab[1] * mv[0]
    qx, qy = px - a[0], py - a[1]
    t = (qx * mv[1] - qy * mv[0]) / det
    u = -(ab[0] * qy - ab[1] * qx) / det
    return t, u * sweep, u


def test_knife_levels_to_its_cap_and_accounts_every_scrap():
    h0 = np.float32(2e-4)
    o = film(flat_state(h0))
    before = o.buf("h").copy()
    a, b, dirv, sweep, height = (20 * DX, 50 * DX), (70 * DX, 50 * DX), (0.0, 1.0), 0.004, 6e-5
    o.set_knife(knife(a, b, dirv, sweep, height))
    o.knife_pass()
    t, behind, u = blade_cells(a, b, dirv, sweep)
    margin = 1e-6
    surely_in = (t > margin) & (t < 1 - margin) & (behind > -DX + 1e-9) & (behind < 0.07 - 1e-9)
    surely_out = (t < -margin) | (t > 1 + margin) | (behind < -DX - 1e-9) | (behind > 0.07 + 1e-9)
    hn = o.buf("h")
    cap = height + np.maximum(behind - 0.01, 0.0) * 0.25
    assert np.allclose(hn[surely_in], np.minimum(cap[surely_in], h0), rtol=1e-5, atol=1e-12)
    assert np.array_equal(hn[surely_out], before[surely_out])
    cut = surely_in & (cap < h0 * (1 - 1e-5))  # the cells the blade takes paint from
    assert cut.sum() > 500
    assert np.all(o.buf("lay")[cut] & 0xFF == 5)
    assert np.all(o.buf("lay")[cut] >> 8 == turn16(0.0 + 1e-9, sweep))
    assert np.array_equal(hn[surely_in & (cap > h0 * (1 + 1e-5))], before[surely_in & (cap > h0 * (1 + 1e-5))])
    sc = o.take_scraps()
    assert len(sc) > 0 and np.all(sc["frame"] == 5)
    key = sc["cell"].astype(np.int64) * 2 + (sc["vol"] < 0)
    assert np.all(np.diff(key) > 0), "ascending cells, taken paint before laid paint"
    lost = np.sum((before - hn).astype(np.float64)) * DX * DX
    assert abs(float(np.sum(sc["vol"].astype(np.float64))) - lost) < 1e-6 * abs(lost)
    assert set(np.unique(sc["cell"]).tolist()) <= set(np.nonzero((before != hn).reshape(-1))[0].tolist())
    assert len(o.take_scraps()) == 0


def test_knife_fill_covers_ratio_to_the_power_0_6():
    global W, H
    saved = (W, H)
    W, H = 512, 400
    try:
        s = flat_state(np.float32(0.0))
        s["strata_pig"][...] = 0
        o = film(s)
        a, b, dirv, sweep, height = (10 * DX, 200 * DX), (500 * DX, 200 * DX), (0.0, 1.0), 0.08, 5e-5
        o.set_knife(knife(a, b, dirv, sweep, height, fill=0.5))
        o.knife_pass()
        t, behind, u = blade_cells(a, b, dirv, sweep)
        edge = (t > 0.001) & (t < 0.999) & (behind > 0) & (behind < 0.07) & (u < 0.999)
        laid = o.buf("h")[edge] > 0
        frac = float(laid.mean())
        assert abs(frac - 0.5 ** 0.6) < 0.025, frac
        assert np.allclose(o.buf("h")[edge][laid], 0.75 * height, rtol=1e-6)
        sc = o.take_scraps()
        assert np.all(sc["vol"] < 0) and np.all(sc["pig"] == 0)
        assert len(sc) == int((o.buf("h") > 0).sum())
    finally:
        W, H = saved
