"""Frames replayed from captured hipGraphs equal directly launched frames
bit for bit: the same scene runs on two films, one with graph replay on, with
host reads, stats polls and deposit uploads between the replays."""
import numpy as np
import pytest

import canvaz_film as czf
import scenes

pytestmark = pytest.mark.gpu

PARTS = ("h", "strata_h", "strata_pig", "strata_cure", "rheo", "lay", "delta", "tile_flags", "tile_moved")


def assert_same(a, b, tag):
    for name in PARTS:
        x, y = a.buffer(name), b.buffer(name)
        assert np.array_equal(x.view(np.uint32), y.view(np.uint32)), f"{tag}: {name} differs"


def test_graph_replay_equals_direct(gpu):
    w = h = 96
    s = scenes.blobs_settings(w, tilt_deg=25.0)
    st = scenes.blobs_state(w, h, seed=5, nblobs=8, sigma_cells=(2.0, 8.0))
    direct = czf.Film(w, h, s)
    graph = czf.Film(w, h, s)
    direct.use_graphs(False)
    graph.use_graphs(True)
    for f in (direct, graph):
        f.flux_shape(0)  # no tuner frames: every repeated shape is replayed
    assert direct.write_state_arrays(st) and graph.write_state_arrays(st)
    pig = scenes.jet_paint()
    for f in range(48):
        deps = []
        if f % 5 == 2:
            x = (0.2 + 0.02 * f) * w * direct.dx
            deps = [czf.deposit_capsule((x, 0.5 * h * direct.dx), (x + 4 * direct.dx, 0.5 * h * direct.dx),
                                        3 * direct.dx, 2e-9, pig, direct.dx)]
        direct.step(deps)
        graph.step(deps)
        if f % 3 == 0:
            sd, sg = direct.stats(), graph.stats()
            assert (sd.h_max, sd.active_tiles, sd.substeps_last) == (sg.h_max, sg.active_tiles, sg.substeps_last)
        if f % 8 == 7:
            assert_same(direct, graph, f"frame {f + 1}")
    direct.sync()
    graph.sync()
    assert_same(direct, graph, "end")
    caps, reps = graph.graph_counts()
    assert caps > 0 and reps > caps, (caps, reps)
    assert direct.graph_counts() == (0, 0)


@pytest.mark.parametrize("batched", [False, True])
def test_flux_shapes_agree(gpu, batched):
    """The launch shapes of the flux and apply passes (window, face-parallel,
    flux + apply in one launch) and the online choice between them give
    bit-identical frames, for a single film and in a batch."""
    w = h = 112
    s = scenes.blobs_settings(w, tilt_deg=35.0)
    st = scenes.blobs_state(w, h, seed=8, nblobs=10, sigma_cells=(2.0, 10.0))
    shapes = (0, 1, 2, -1)
    films = []
    for shape in shapes:
        f = czf.Film(w, h, s)
        assert f.write_state_arrays(st)
        films.append(f)
        if not batched:
            f.flux_shape(shape)
    batches = []
    if batched:
        for f, shape in zip(films, shapes):
            b = czf.Batch(2)
            b.flux_shape(shape)
            b.add(f)
            batches.append(b)
    for _ in range(16):
        if batched:
            for b in batches:
                b.step()
        else:
            for f in films:
                f.step()
    for b in batches:
        b.sync()
    for f in films:
        f.sync()
    for k, shape in enumerate(shapes[1:], 1):
        assert_same(films[0], films[k], f"window vs shape {shape}")
    if not batched:
        used = films[2].stats().flux_shape_cell_substeps
        assert used[2] > 0 and used[0] == used[1] == 0, list(used)
