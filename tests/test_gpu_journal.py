"""The device undo journal (k_journal_mark / k_journal_list / k_journal_save
and the restore kernels; shaders/journal.wgsl, src/film/mod.rs:357-443,
1246-1275) against the restated journal: after every frame the ring head,
the slot records, the per-tile marks and the saved cell words of every live
record are bit-identical, and so is the state after each restore; frames
replayed from hipGraphs and direct ones alike.  On a larger canvas the
session's undo (UndoHistory) brings the state back bit for bit."""
import numpy as np
import pytest

import canvaz_film as czf
import cz_types
import scenes
from helpers import make_pair, oracle_capsule

pytestmark = pytest.mark.gpu

PARTS = ("h", "m", "rheo", "strata_h", "strata_pig", "strata_cure", "sub_refl", "lay")
TILE_BYTES = 53 * 256 * 4


def _aos_index():
    """Word index in a device record of word w of cell li: the record holds
    the parts one after another (first word w0, n words per cell), each as
    its 256 cells' words."""
    parts = [(0, 1), (1, 1), (2, 2), (4, 4), (8, 32), (40, 4), (44, 8), (52, 1)]
    idx = np.zeros((256, 53), np.int64)
    for w0, n in parts:
        for k in range(n):
            idx[:, w0 + k] = 256 * w0 + np.arange(256) * n + k
    return idx


AOS_INDEX = _aos_index()


def compare_journal(g, o, w, h, where):
    head, cap = g.journal_head()
    assert head == o.journal_head(), f"{where}: head {head} vs {o.journal_head()}"
    mg = g.buffer("jmeta").reshape(-1, 4)
    mo = o.buf("jmeta").reshape(-1, 4)
    assert mg.shape == mo.shape == (cap, 4)
    bad = np.argwhere(mg != mo)
    assert bad.size == 0, f"{where}: jmeta differs at slots {bad[:6, 0].tolist()}: {mg[bad[0][0]]} vs {mo[bad[0][0]]}"
    assert np.array_equal(g.buffer("jmark"), o.buf("jmark")), f"{where}: jmark"
    dg = g.buffer("jdata").reshape(cap, 256 * 53)[:, AOS_INDEX]       # to [slot][cell][word]
    do = o.buf("jdata").reshape(cap, 256, 53)
    tw = (w + 15) // 16
    li = np.arange(256)
    for s in np.nonzero(mo[:, 3] == 1)[0]:
        t = int(mo[s, 0])
        inb = ((t % tw) * 16 + (li & 15) < w) & ((t // tw) * 16 + (li >> 4) < h)
        a, b = dg[s][inb], do[s][inb]
        if not np.array_equal(a, b):
            c, word = np.argwhere(a != b)[0]
            pytest.fail(f"{where}: record {s} (tile {t}) cell {c} word {word}: {a[c, word]:#x} vs {b[c, word]:#x}")


def compare_state(g, o, where):
    for k in PARTS:
        a, b = g.buffer(k).view(np.uint32), o.buf(k).view(np.uint32)
        bad = np.argwhere(a != b)
        assert bad.size == 0, f"{where}: {k} differs at {bad[:4].tolist()}"
    assert np.array_equal(g.buffer("tile_moved"), o.buf("tile_moved")), f"{where}: tile_moved"


def capsule_pair(g, p0, p1, r, vol, pig):
    return [czf.deposit_capsule(p0, p1, r, vol, pig, g.dx)], [oracle_capsule(p0, p1, r, vol, pig, g.dx)]


@pytest.mark.parametrize("graphs", [True, False])
def test_journal_and_restore_equal_restatement(gpu, graphs):
    w, h = 120, 88  # partial edge tiles
    s = scenes.blobs_settings(w, tilt_deg=30.0)
    s.height_m = h * s.width_m / w
    st = scenes.blobs_state(w, h, seed=21, nblobs=6, sigma_cells=(2.0, 8.0))
    g, o = make_pair(w, h, s, st)
    g.use_graphs(graphs)
    g.journal_budget(256 * TILE_BYTES)
    assert g.journal_head() == (0, 256)
    pig = scenes.jet_paint()
    dx = g.dx
    plan = {3: "gen1", 9: "gen2"}
    for f in range(20):
        if f in plan:
            gen = 1 if f == 3 else 2
            g.journal_gen(gen)
            o.journal_begin(gen, 256)
        dg, do = [], []
        if f in (3, 4, 9):
            y = (0.3 if f < 9 else 0.7) * h * dx
            dg, do = capsule_pair(g, (0.2 * w * dx + f * dx, y), (0.6 * w * dx, y + 5 * dx), 3 * dx, 2e-8, pig)
        if f == 6:
            for film in (g, o):
                film.touch_rect(0.8 * w * dx, 0.0, 2.0 * w * dx, 0.2 * h * dx)
        if f == 15:
            g.restore_journal(2)
            o.restore_journal(2)
            compare_state(g, o, "after restore(2)")
        g.step(dg)
        o.step(do)
        g.sync()
        if f >= 3:
            compare_journal(g, o, w, h, f"frame {f}")
    compare_state(g, o, "frame 19")
    g.restore_journal(1)
    o.restore_journal(1)
    compare_state(g, o, "after restore(1)")
    if graphs:
        assert g.graph_counts()[1] > 0, "some frames should replay from a graph"


def test_ring_overflow_in_one_frame_equals_restatement(gpu):
    w, h = 256, 184  # 16 x 12 = 192 tiles against a ring of 64
    s = scenes.blobs_settings(w, tilt_deg=20.0)
    s.height_m = h * s.width_m / w
    st = scenes.blobs_state(w, h, seed=8, nblobs=20, sigma_cells=(2.0, 10.0))
    g, o = make_pair(w, h, s, st)
    g.journal_budget(0)
    for f in range(2):
        g.step([])
        o.step([])
    g.journal_gen(1)
    o.journal_begin(1, 64)
    for film in (g, o):
        film.touch_rect(-1.0, -1.0, 1.0, 1.0)
    for f in range(3):
        g.step([])
        o.step([])
        g.sync()
        compare_journal(g, o, w, h, f"frame {f}")
    assert g.journal_head()[0] == 192
    g.journal_gen(2)
    o.journal_begin(2, 64)
    for f in range(3):
        g.step([])
        o.step([])
        g.sync()
        compare_journal(g, o, w, h, f"gen 2 frame {f}")
    for target in (2, 1):
        g.restore_journal(target)
        o.restore_journal(target)
        compare_state(g, o, f"restore({target})")


def differing(a, b):
    """Names of the state parts that differ (read_state lists)."""
    return [PARTS[k] for k in range(8) if a[k] != b[k]]


def test_undo_history_round_trip_at_512(gpu):
    w, h = 512, 384
    s = scenes.blobs_settings(w, tilt_deg=25.0)
    s.height_m = h * s.width_m / w
    # Curing goes on in tiles no frame journals (paint at rest), and restore
    # leaves them as they are, as the reference's does; without curing every
    # change is journaled and undo is exact.
    s.cure_rate = 0.0
    g = czf.Film(w, h, s)
    g.journal_budget(4096 * TILE_BYTES)  # room for every tile of the three levels
    for _ in range(3):
        g.step([])
    u = czf.UndoHistory(g)
    pig = scenes.jet_paint()
    dx = g.dx

    def paint(y_frac, frames=8):
        for f in range(frames):
            deps = []
            if f < 3:
                x = (0.1 + 0.25 * f) * w * dx
                deps = [czf.deposit_capsule((x, y_frac * h * dx), (x + 0.15 * w * dx, (y_frac + 0.1) * h * dx),
                                            6 * dx, 3e-8, pig, dx)]
            g.step(deps)
        g.sync()
        print(f"painted at {y_frac}: head {g.journal_head()[0]}", flush=True)

    s0 = g.read_state()
    u.begin_stroke()
    paint(0.3)
    s1 = g.read_state()
    assert differing(s1, s0)
    u.begin_stroke()
    paint(0.6)
    assert differing(g.read_state(), s1)
    u.rollback()
    assert differing(g.read_state(), s1) == []
    u.begin_full()
    paint(0.45)
    u.rollback()
    assert differing(g.read_state(), s1) == []
    u.rollback()
    assert differing(g.read_state(), s0) == []
    assert u.gen == 6 and g.journal_head()[1] == 4096


def test_undo_refused_when_the_ring_overflowed(gpu):
    w, h = 256, 256
    s = scenes.blobs_settings(w)
    s.height_m = s.width_m
    g = czf.Film(w, h, s)
    g.journal_budget(0)  # 64 tiles
    u = czf.UndoHistory(g)
    u.begin_stroke()
    g.touch_rect(-1.0, -1.0, 1.0, 1.0)  # all 256 tiles
    g.step([])
    g.sync()
    with pytest.raises(czf.FilmError, match="changed too much"):
        u.rollback()


def test_batch_refuses_a_journaling_film(gpu):
    s = scenes.blobs_settings(64)
    g = czf.Film(64, 64, s)
    assert g.journal_head() == (0, 7419)  # the default 384 MiB ring
    g.journal_gen(1)
    b = czf.Batch(4)
    with pytest.raises(czf.FilmError, match="undo journal"):
        b.add(g)
    g.journal_gen(0)
    k = cz_types.Knife()
    k.a[:] = (0.0, 0.0)
    k.b[:] = (0.01, 0.0)
    k.dir[:] = (0.0, 1.0)
    k.sweep, k.active = 0.001, 1
    g.set_knife(k)
    with pytest.raises(czf.FilmError, match="knife"):
        b.add(g)
    g.set_knife(None)
    b.add(g)
    with pytest.raises(czf.FilmError, match="batch"):
        g.journal_gen(2)
    b.remove(g)
