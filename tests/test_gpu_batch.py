"""Batched film step: every member of a batch ends each frame bit-identical
to the same film stepped alone (the stated bar is 1e-5 relative; the batched
launches run the same per-cell arithmetic, so the test asks for equality).
Members differ in size, tilt, substep count and implicit steps, and join and
leave between frames."""
import numpy as np
import pytest

import canvaz_film as czf
import scenes

pytestmark = pytest.mark.gpu

PARTS = ("h", "strata_h", "strata_pig", "strata_cure", "rheo", "lay", "delta", "fnorm", "tile_flags",
         "tile_moved", "m", "sub_refl")


def assert_same(a, b, tag):
    for name in PARTS:
        x, y = a.buffer(name), b.buffer(name)
        if not np.array_equal(x.view(np.uint32), y.view(np.uint32)):
            bad = np.flatnonzero(x.view(np.uint32).ravel() != y.view(np.uint32).ravel())
            raise AssertionError(f"{tag}: {name} differs at {bad.size} words, first {bad[:6]}")
    sa, sb = a.stats(), b.stats()
    assert (sa.h_max, sa.mu_min, sa.active_tiles, sa.substeps_last, sa.implicit_steps) == \
           (sb.h_max, sb.mu_min, sb.active_tiles, sb.substeps_last, sb.implicit_steps), tag
    assert list(sa.cg_report) == list(sb.cg_report), tag


def scene(k):
    """Canvas k of the mixed test batch: (w, h, settings, state)."""
    specs = [
        (64, 64, 15.0, dict(seed=1, nblobs=6, sigma_cells=(2.0, 8.0), peak=(20e-6, 60e-6),
                            mu=(20.0, 50.0))),  # thin viscous paint: few substeps
        (96, 80, 40.0, dict(seed=2, nblobs=8, sigma_cells=(2.0, 10.0))),
        (96, 96, 5.0, dict(seed=5, nblobs=6, sigma_cells=(6.0, 12.0), peak=(4e-3, 6e-3), mu=(0.05, 0.1),
                           tau=(1e-3, 2e-3))),  # thick fluid paint: implicit steps
        (48, 112, 60.0, dict(seed=4, nblobs=5, sigma_cells=(2.0, 6.0))),
        (128, 128, 25.0, dict(seed=6, nblobs=12, sigma_cells=(2.0, 14.0))),
    ]
    w, h, tilt, kw = specs[k]
    return w, h, scenes.blobs_settings(w, tilt_deg=tilt), scenes.blobs_state(w, h, **kw)


def make(k):
    w, h, s, st = scene(k)
    f = czf.Film(w, h, s)
    assert f.write_state_arrays(st)
    return f


def test_batch_equals_alone(gpu):
    n = 5
    alone = [make(k) for k in range(n)]
    inb = [make(k) for k in range(n)]
    batch = czf.Batch(8)
    for f in inb[:4]:
        batch.add(f)
    implicit_seen = False
    for frame in range(14):
        if frame == 3:
            batch.add(inb[4])          # joins late: both twins had stood still so far
        if frame == 7:
            batch.remove(inb[1])       # leaves: continues on its own
        for k, f in enumerate(alone):
            if k == 4 and frame < 3:
                continue
            f.step()
        if frame >= 7:
            inb[1].step()
        batch.step()
        if frame % 4 == 3:
            batch.sync()
            for k in range(n):
                alone[k].sync()
                inb[k].sync()
                assert_same(alone[k], inb[k], f"member {k} frame {frame + 1}")
                implicit_seen |= inb[k].stats().implicit_steps > 0
    batch.sync()
    for k in range(n):
        alone[k].sync()
        inb[k].sync()
        assert_same(alone[k], inb[k], f"member {k} end")
    subs = {inb[k].stats().substeps_last for k in range(n)}
    assert len(subs) > 1, "members should differ in substep count"
    assert implicit_seen, "no member exercised the implicit solve"


def test_batch_graphs_equal_direct(gpu):
    direct = [make(k) for k in (0, 1, 3)]
    graph = [make(k) for k in (0, 1, 3)]
    bd, bg = czf.Batch(4), czf.Batch(4)
    bd.use_graphs(False)
    bg.use_graphs(True)
    for b in (bd, bg):
        b.flux_shape(0)  # no tuner frames: every repeated shape is replayed
    for f in direct:
        bd.add(f)
    for f in graph:
        bg.add(f)
    for frame in range(24):
        bd.step()
        bg.step()
        if frame % 3 == 2:
            bd.sync()
            bg.sync()
            for a, b in zip(direct, graph):
                assert_same(a, b, f"graph frame {frame + 1}")
    bd.sync()
    bg.sync()
    for a, b in zip(direct, graph):
        assert_same(a, b, "graph end")
    caps, reps = bg.graph_counts()
    assert caps > 0 and reps > 0, (caps, reps)
    assert bd.graph_counts() == (0, 0)


def test_member_calls_are_refused(gpu):
    f = make(0)
    b = czf.Batch(2)
    b.add(f)
    with pytest.raises(czf.FilmError):
        f.step()
    b.step()
    b.sync()
    assert f.stats().frame == 1
    b.remove(f)
    f.step()
    f.sync()
    assert f.stats().frame == 2


@pytest.mark.parametrize("shape", [3, 4])
def test_batch_wide_shapes_equal_alone(gpu, shape):
    """The 1024-thread substep shapes in a batch: members equal to the same
    films stepped alone in the default shapes."""
    alone = [make(k) for k in (1, 2, 4)]
    inb = [make(k) for k in (1, 2, 4)]
    b = czf.Batch(4)
    b.flux_shape(shape)
    for f in inb:
        b.add(f)
    for frame in range(8):
        b.step()
        for f in alone:
            f.step()
    b.sync()
    for k in range(3):
        alone[k].sync()
        assert_same(alone[k], inb[k], f"shape {shape} member {k}")


def test_batch_cell_index_equals_alone(gpu):
    """Members with an index per cell (settings.n_per_cell) beside one
    without: each equal to its twin stepped alone, cell indices included."""
    def make_n(k, per_cell):
        w, h, s, st = scene(k)
        if per_cell:
            s.n_per_cell = 1
            st = scenes.blobs_state(w, h, **dict(seed=k + 1, nblobs=8, sigma_cells=(2.0, 10.0),
                                                  n_range=(0.4, 1.0)))
        f = czf.Film(w, h, s)
        assert f.write_state_arrays(st)
        return f
    plan = [(1, True), (3, False), (4, True)]
    alone = [make_n(k, p) for k, p in plan]
    inb = [make_n(k, p) for k, p in plan]
    batch = czf.Batch(4)
    for f in inb:
        batch.add(f)
    for frame in range(8):
        for f in alone:
            f.step()
        batch.step()
    batch.sync()
    for i, (k, p) in enumerate(plan):
        alone[i].sync()
        assert_same(alone[i], inb[i], f"member {k}")
        if p:
            a, b = alone[i].buffer("n_cell"), inb[i].buffer("n_cell")
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), f"member {k}: cell indices differ"
            wet = alone[i].buffer("h") > 0
            assert a[wet].min() < 0.6 and a[wet].max() > 0.8
