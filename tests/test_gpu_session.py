"""A painting session end to end against the restated film: 40 frames of one
canvas with every per-frame input the reference's App feeds the film
(src/main.rs:1875-1909, 2436-2444): deposits, brush contacts of every kind,
a knife over several frames, journal generations begun per stroke, an undo
in the middle, all through graph-replayed frames.  After every frame the
state, the moved tiles, the ring head and the scraps taken are bit-identical,
and so is the reservoir at the end."""
import math

import numpy as np
import pytest

import canvaz_film as czf
import cz_types
import scenes
from helpers import make_pair, oracle_capsule

pytestmark = pytest.mark.gpu

PARTS = ("h", "m", "rheo", "strata_h", "strata_pig", "strata_cure", "sub_refl", "lay")


@pytest.mark.parametrize("w,h", [(144, 112), (75, 53)])
def test_session_of_strokes_tools_and_undo(gpu, w, h):
    # 75 x 53: partial tiles on both axes and rows that are not 16-byte aligned
    s = scenes.blobs_settings(w, tilt_deg=25.0)
    s.height_m = h * s.width_m / w
    st = scenes.blobs_state(w, h, seed=77, nblobs=7, sigma_cells=(3.0, 9.0))
    g, o = make_pair(w, h, s, st)
    g.journal_budget(512 * 53 * 256 * 4)
    rng = np.random.default_rng(12)
    dx = g.dx
    pig = scenes.jet_paint()
    gen = 0
    for f in range(40):
        if f % 8 == 0:  # a new stroke: a new journal generation (begin_stroke)
            gen += 1
            g.journal_gen(gen)
            o.journal_begin(gen, 512)
        if f == 27:  # undo the stroke begun at frame 24 (rollback), then a new generation
            g.restore_journal(gen)
            o.restore_journal(gen)
            gen += 1
            g.journal_gen(gen)
            o.journal_begin(gen, 512)
        dg, do = [], []
        if f % 3 == 0:
            x, y = rng.uniform(0.2, 0.8) * w * dx, rng.uniform(0.2, 0.8) * h * dx
            p1 = (x + 10 * dx, y + rng.uniform(-5, 5) * dx)
            dg = [czf.deposit_capsule((x, y), p1, 3 * dx, 1.5e-8, pig, dx)]
            do = [oracle_capsule((x, y), p1, 3 * dx, 1.5e-8, pig, dx)]
        if f % 4 == 1:
            cs = []
            for q in range(30):
                c = cz_types.Contact()
                x, y = rng.uniform(0.1, 0.9) * w * dx, rng.uniform(0.1, 0.9) * h * dx
                c.p0[:] = (x, y)
                c.p1[:] = (x + rng.uniform(-4, 4) * dx, y + rng.uniform(-4, 4) * dx)
                c.r, c.zb = rng.uniform(1.5, 4.0) * dx, rng.uniform(0.0, 2e-4)
                c.bristle, c.pad = int(rng.integers(0, 40)), int(q % 4)
                cs.append(c)
            g.upload_contacts(cs)
            o.upload_contacts(cs)
        if 10 <= f < 16:
            k = cz_types.Knife()
            a = (0.25 * w * dx, (0.3 + 0.02 * (f - 10)) * h * dx)
            k.a[:] = a
            k.b[:] = (a[0] + 0.4 * w * dx, a[1] + 0.05 * h * dx)
            ang = math.atan2(0.05 * h, 0.4 * w) + math.pi / 2
            k.dir[:] = (math.cos(ang), math.sin(ang))
            k.sweep, k.height, k.half_width, k.active, k.frame, k.bands = 3 * dx, 6e-5, 0.01, 1, f, 5
            k.fill[:] = list(rng.uniform(0, 1, 8))
            k.smear[:] = list(rng.uniform(0, 0.5, 8))
            g.set_knife(k)
            o.set_knife(k)
        if f == 16:
            g.set_knife(None)
            o.set_knife(None)
        g.step(dg)
        o.step(do)
        g.sync()
        for name in PARTS + ("tile_moved",):
            a_, b_ = g.buffer(name).view(np.uint32), o.buf(name).view(np.uint32)
            bad = np.argwhere(a_ != b_)
            assert bad.size == 0, f"frame {f}: {name} differs at {bad[:4].tolist()}"
        assert g.journal_head()[0] == o.journal_head(), f"frame {f}: ring head"
        sg, so = g.take_scraps(), o.take_scraps()
        assert np.array_equal(sg.view(np.uint32).reshape(-1, 12), so.view(np.uint32).reshape(-1, 12)), f"frame {f}"
    assert np.array_equal(g.take_reservoir(), o.take_reservoir())
    assert g.graph_counts()[1] > 0
