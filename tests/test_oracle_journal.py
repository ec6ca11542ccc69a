"""The restated undo journal (shaders/journal.wgsl, src/film/mod.rs:357-443,
1246-1275) pinned by what its rules fix independently of its code: the tile
set of touch_rect from the rectangle's closed form, the busy set as a 3x3
dilation of the last frame's active tiles (computed here with numpy), one
record per tile per generation in ascending tile order, the ring keeping
the newest cap records, and restore bringing back, bit for bit, the state a
generation started from on a canvas where every change comes from a
journaled tile."""
import numpy as np
import pytest

import scenes
from helpers import make_pair, oracle_capsule

PARTS = ("h", "m", "rheo", "strata_h", "strata_pig", "strata_cure", "sub_refl", "lay")


def quiet_film(w, h, state=None, frames=4):
    s = scenes.blobs_settings(w, tilt_deg=30.0)
    s.height_m = h * s.width_m / w
    _, o = make_pair(w, h, s, state, gpu=False)
    for _ in range(frames):
        o.step([])
    return o, s


def records(o):
    """(tile, gen, seq) of the valid slots, by sequence number."""
    m = o.buf("jmeta").reshape(-1, 4)
    v = m[m[:, 3] == 1]
    return sorted((int(t), int(g), int(q)) for t, g, q, _ in v), m


def snapshot(o):
    return {k: o.buf(k).copy() for k in PARTS}


def diff_cells(o, snap):
    return {k: int(np.count_nonzero(o.buf(k).view(np.uint32) != snap[k].view(np.uint32))) for k in PARTS}


def test_touch_rect_tiles_follow_the_closed_form():
    w, h = 80, 48  # 5 x 3 tiles of 16 cells
    o, s = quiet_film(w, h)
    assert not np.any(o.buf("tile_flags") & 32), "the empty canvas must be at rest"
    tm = 16 * s.width_m / w
    cases = [
        ((0.5 * tm, 0.5 * tm, 1.5 * tm, 0.2 * tm), [0, 1]),
        ((2.2 * tm, 1.1 * tm, 2.3 * tm, 2.9 * tm), [7, 12]),
        ((-5.0, -5.0, -1e-9, 1.0), []),                 # max x below 0: nothing
        ((float("nan"), 0.0, 1.0, 1.0), []),            # not finite: nothing
        ((-1e30, -1e30, 1e30, 1e30), list(range(15))),  # clamped to the canvas
        ((4.99 * tm, 2.5 * tm, 9.0 * tm, 9.0 * tm), [14]),
        ((-3.0 * tm, -3.0 * tm, 0.1 * tm, 0.1 * tm), [0]),
    ]
    head = 0
    for gen, (rect, tiles) in enumerate(cases, start=1):
        o.journal_begin(gen, 64)
        o.touch_rect(*rect)
        o.step([])
        assert o.journal_head() == head + len(tiles), rect
        got = [r for r in records(o)[0] if r[1] == gen]
        assert got == [(t, gen, head + k) for k, t in enumerate(tiles)], rect
        assert sorted(np.nonzero(o.buf("jmark") == gen)[0].tolist()) == tiles
        head += len(tiles)
        # the same touch again in the same generation saves nothing
        o.touch_rect(*rect)
        o.step([])
        assert o.journal_head() == head


def test_generation_zero_records_nothing():
    o, s = quiet_film(48, 48)
    o.touch_rect(0.0, 0.0, 1.0, 1.0)  # ignored while the journal is off
    o.journal_begin(0, 64)
    o.touch_rect(0.0, 0.0, 1.0, 1.0)
    o.step([])
    assert o.journal_head() == 0
    assert not np.any(o.buf("jmeta"))


def test_busy_tiles_are_the_dilated_active_set():
    w, h = 128, 96
    st = scenes.blobs_state(w, h, seed=5, nblobs=3, sigma_cells=(3.0, 6.0))
    o, s = quiet_film(w, h, st, frames=3)
    tw, th = w // 16, h // 16
    head = 0
    for gen in (1, 2, 3):
        act = (o.buf("tile_flags").reshape(th, tw) & 32) != 0
        assert act.any() and not act.all()
        pad = np.pad(act, 1)
        busy = np.zeros_like(act)
        for dy in (0, 1, 2):
            for dx in (0, 1, 2):
                busy |= pad[dy:dy + th, dx:dx + tw]
        want = np.nonzero(busy.reshape(-1))[0].tolist()
        o.journal_begin(gen, 256)
        o.step([])
        got = [r for r in records(o)[0] if r[1] == gen]
        assert got == [(t, gen, head + k) for k, t in enumerate(want)]
        head += len(want)
        assert o.journal_head() == head


def test_ring_keeps_the_newest_records():
    w = h = 192  # 144 tiles, ring of 64
    o, s = quiet_film(w, h)
    o.journal_begin(1, 64)
    o.touch_rect(-1.0, -1.0, 1e9, 1e9)
    o.step([])
    assert o.journal_head() == 144
    recs, m = records(o)
    assert recs == [(q, 1, q) for q in range(80, 144)]  # ascending tiles: tile = sequence number
    for t, g, q in recs:
        assert tuple(m[q % 64][:3]) == (t, g, q)
    # restore reaches only the tiles whose records survived
    o.buf("tile_moved")[...] = 0
    o.restore_journal(1)
    assert np.nonzero(o.buf("tile_moved"))[0].tolist() == list(range(80, 144))


def stroke(o, s, w, h, frames, y_frac, pig, head_frames=3):
    dx = s.width_m / w
    for f in range(frames):
        deps = []
        if f < head_frames:
            x0 = (0.15 + 0.2 * f) * w * dx
            p0, p1 = (x0, y_frac * h * dx), (x0 + 0.15 * w * dx, (y_frac + 0.05) * h * dx)
            deps = [oracle_capsule(p0, p1, 4 * dx, 6e-8, pig, dx)]
        o.step(deps)


def test_restore_returns_each_generations_start():
    w, h = 128, 80
    o, s = quiet_film(w, h)
    pig = scenes.jet_paint()
    s0 = snapshot(o)
    o.journal_begin(1, 256)  # both generations fit the ring (40 tiles each at most)
    stroke(o, s, w, h, 12, 0.3, pig)
    s1 = snapshot(o)
    assert sum(diff_cells(o, s0).values()) > 2000
    o.journal_begin(2, 256)
    stroke(o, s, w, h, 12, 0.55, pig)
    assert sum(diff_cells(o, s1).values()) > 2000
    o.restore_journal(2)
    assert diff_cells(o, s1) == dict.fromkeys(PARTS, 0)
    o.restore_journal(1)
    assert diff_cells(o, s0) == dict.fromkeys(PARTS, 0)


def test_restore_leaves_unjournaled_curing_alone():
    """Paint at rest keeps curing in tiles the generation has not journaled
    yet (a tile is saved when a frame first touches it or finds it busy);
    restore brings a tile back to the frame it was saved in, or leaves it
    (journal.wgsl:109-123).  So the heights, strata and lay come back exactly
    and only cure levels (and the rheology that follows them) differ."""
    w, h = 192, 144
    s = scenes.blobs_settings(w, tilt_deg=25.0)
    s.height_m = h * s.width_m / w
    _, o = make_pair(w, h, s, None, gpu=False)
    for _ in range(3):
        o.step([])
    pig = scenes.jet_paint()
    dx = s.width_m / w

    def paint(yf):
        for f in range(8):
            deps = []
            if f < 3:
                x = (0.1 + 0.25 * f) * w * dx
                deps = [oracle_capsule((x, yf * h * dx), (x + 0.15 * w * dx, (yf + 0.1) * h * dx), 6 * dx, 4.2e-9,
                                       pig, dx)]
            o.step(deps)

    o.journal_begin(1, 4096)
    paint(0.3)
    s1 = snapshot(o)
    o.journal_begin(2, 4096)
    paint(0.6)
    o.restore_journal(2)
    d = diff_cells(o, s1)
    assert d["h"] == 0 and d["strata_h"] == 0 and d["strata_pig"] == 0 and d["lay"] == 0
    assert d["strata_cure"] > 0


@pytest.mark.parametrize("target", [1, 2, 3])
def test_restore_takes_the_oldest_record_at_or_after_target(target):
    """Three generations each touching tile 0 with a different state: the
    restore brings back the state saved by generation `target`."""
    w = h = 32
    o, s = quiet_film(w, h)
    saved = {}
    for gen in (1, 2, 3):
        o.buf("h")[:16, :16] = 1e-6 * gen  # a distinct state per generation (written between frames)
        o.journal_begin(gen, 64)
        saved[gen] = o.buf("h").copy()
        o.touch_rect(0.0, 0.0, 1e-9, 1e-9)
        o.step([])
    o.buf("h")[...] = 7e-6
    o.restore_journal(target)
    np.testing.assert_array_equal(o.buf("h")[:16, :16], saved[target][:16, :16])
    assert np.all(o.buf("h")[:16, 16:] == 7e-6)  # tile 1 has no record
