"""The device tool passes (k_brush; k_knife_count / k_knife_scan /
k_knife_apply; the contact binning of upload_contacts) against the restated
ones over whole frames: every state part, the moved tiles, the per-bristle
reservoir and the knife's scrap list are bit-identical, for every contact
kind, knives at several angles (the nearly degenerate blade included, which
visits the whole canvas), with fill and smear, frames replayed from
hipGraphs and direct."""
import math

import numpy as np
import pytest

import cz_types
import scenes
from helpers import make_pair

pytestmark = pytest.mark.gpu

PARTS = ("h", "m", "rheo", "strata_h", "strata_pig", "strata_cure", "sub_refl", "lay")


def compare(g, o, where):
    for k in PARTS + ("tile_moved",):
        a, b = g.buffer(k).view(np.uint32), o.buf(k).view(np.uint32)
        bad = np.argwhere(a != b)
        assert bad.size == 0, f"{where}: {k} differs at {bad[:4].tolist()}"


def setup(w, h, seed):
    s = scenes.blobs_settings(w, tilt_deg=20.0)
    s.height_m = h * s.width_m / w
    st = scenes.blobs_state(w, h, seed=seed, nblobs=12, sigma_cells=(4.0, 14.0), peak=(5e-5, 6e-4))
    return make_pair(w, h, s, st)


def contacts(rng, w, h, dx, n):
    out = []
    for q in range(n):
        c = cz_types.Contact()
        x, y = rng.uniform(0.1, 0.9) * w * dx, rng.uniform(0.1, 0.9) * h * dx
        ang = rng.uniform(0, 2 * math.pi)
        ln = rng.uniform(0, 12) * dx
        c.p0[:] = (x, y)
        c.p1[:] = (x + ln * math.cos(ang), y + ln * math.sin(ang))
        c.r = rng.uniform(1.5, 6.0) * dx
        c.zb = rng.uniform(-1e-4, 3e-4)
        c.bristle = int(rng.integers(0, 520))  # some beyond the 512 reservoirs
        c.pad = int(rng.integers(0, 4))
        out.append(c)
    return out


@pytest.mark.parametrize("graphs", [True, False])
def test_brush_frames_equal_restatement(gpu, graphs):
    w, h = 160, 128
    g, o = setup(w, h, 31)
    g.use_graphs(graphs)
    rng = np.random.default_rng(4)
    dx = g.dx
    for f in range(8):
        cs = contacts(rng, w, h, dx, 400 if f % 3 else 40)
        if f == 5:
            cs = cs[:3]
        g.upload_contacts(cs)
        o.upload_contacts(cs)
        g.step([])
        o.step([])
        g.sync()
        compare(g, o, f"frame {f}")
        if f % 2 == 1:
            rg, ro = g.take_reservoir(), o.take_reservoir()
            assert np.array_equal(rg, ro), f"frame {f}: reservoir"
            assert rg.any()


def knife(rng, w, h, dx, frame, angle, sweep, degenerate=False):
    k = cz_types.Knife()
    cx, cy = rng.uniform(0.3, 0.7) * w * dx, rng.uniform(0.3, 0.7) * h * dx
    half = rng.uniform(10, 30) * dx
    ca, sa = math.cos(angle), math.sin(angle)
    k.a[:] = (cx - half * ca, cy - half * sa)
    k.b[:] = (cx + half * ca, cy + half * sa)
    perp = angle + (1e-4 if degenerate else math.pi / 2 + rng.uniform(-0.4, 0.4))
    k.dir[:] = (math.cos(perp), math.sin(perp))
    k.sweep = sweep
    k.height = rng.uniform(2e-5, 2e-4)
    k.half_width = half
    k.active = 1
    k.frame = frame
    k.bands = int(rng.integers(0, 10))
    for bi in range(8):
        frac = np.zeros(16, np.float32)
        frac[:12] = rng.dirichlet(np.full(12, 0.5))
        frac[13], frac[14] = rng.uniform(0.5, 20), rng.uniform(9, 13)
        k.pig[bi][:] = [int(v) for v in scenes.pack_pig(frac)]
    k.fill[:] = list(rng.uniform(0, 1, 8))
    k.smear[:] = list(rng.uniform(0, 0.6, 8))
    return k


@pytest.mark.parametrize("graphs", [True, False])
def test_knife_frames_equal_restatement(gpu, graphs):
    w, h = 160, 128
    g, o = setup(w, h, 17)
    g.use_graphs(graphs)
    rng = np.random.default_rng(9)
    dx = g.dx
    plan = [(0.0, 8 * dx), (0.0, 8 * dx), (0.7, 12 * dx), (2.5, 5 * dx), (0.3, 0.02),
            (1.2, 10 * dx, True), (4.0, -9 * dx), (0.0, 8 * dx)]
    for f, step in enumerate(plan):
        k = knife(rng, w, h, dx, f + 100, step[0], step[1], degenerate=len(step) > 2)
        if f == 1:
            k = prev  # the same knife twice: a replayed frame shape  # noqa: F821
        g.set_knife(k)
        o.set_knife(k)
        prev = k
        g.step([])
        o.step([])
        g.sync()
        compare(g, o, f"frame {f}")
        sg, so = g.take_scraps(), o.take_scraps()
        assert len(sg) == len(so), f"frame {f}: {len(sg)} vs {len(so)} scraps"
        assert np.array_equal(sg.view(np.uint32).reshape(-1, 12), so.view(np.uint32).reshape(-1, 12)), f"frame {f}"
    g.set_knife(None)
    o.set_knife(None)
    g.step([])
    o.step([])
    g.sync()
    compare(g, o, "knife off")
    assert len(g.take_scraps()) == 0


def test_scraps_accumulate_across_frames_in_frame_order(gpu):
    w, h = 128, 96
    g, o = setup(w, h, 5)
    rng = np.random.default_rng(2)
    dx = g.dx
    for f in range(3):
        k = knife(rng, w, h, dx, f, 0.2 * f, 6 * dx)
        g.set_knife(k)
        o.set_knife(k)
        g.step([])
        o.step([])
    sg, so = g.take_scraps(), o.take_scraps()
    assert len(sg) > 0 and len(sg) == len(so)
    assert np.array_equal(sg.view(np.uint32).reshape(-1, 12), so.view(np.uint32).reshape(-1, 12))
    fr = sg["frame"]
    assert np.all(np.diff(fr.astype(np.int64)) >= 0)
    compare(g, o, "after 3 knife frames")
