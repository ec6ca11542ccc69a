"""The slump pass on the GPU (settings.slump: shed, slump and commit_tmp after
every substep's nose, film.wgsl:563-801, 866-876) against the restatement:
heights, strata and pigments within 1e-5 relative, tile flags, lists, nose
offsets, lay tags and cure masks bit-exact, on canvases whose narrow tall
bumps stand above the cone caps, with the substep shapes tuned and pinned,
and on the jet scene, whose deposits pile up faster than the paint runs."""
import numpy as np
import pytest

import canvaz_film as czf
import scenes
from helpers import make_pair, oracle_capsule
from test_gpu_parity import compare, compare_masks, run_pair

pytestmark = pytest.mark.gpu


def spiky_scene(w, cell_n=False, tilt=20.0):
    s = scenes.blobs_settings(w, tilt_deg=tilt)
    s.slump = 1
    if cell_n:
        s.n_index = 0.7
        s.n_per_cell = 1
    st = scenes.blobs_state(w, w, seed=31, nblobs=16, sigma_cells=(0.8, 2.5), peak=(2e-3, 9e-3),
                            tau=(0.5, 20.0), mu=(1.0, 20.0), n_range=(0.4, 1.0) if cell_n else None)
    return s, st


def test_scene_sheds():
    """The scene exercises the pass: one frame of the restatement with the
    slump on differs from the same frame with it off."""
    w = 96
    s, st = spiky_scene(w)
    _, on = make_pair(w, w, s, st, gpu=False)
    s.slump = 0
    _, off = make_pair(w, w, s, st, gpu=False)
    on.step()
    off.step()
    assert not np.array_equal(on.buf("h"), off.buf("h"))


@pytest.mark.parametrize("shape", [-1, 0, 1, 2, 3, 4, 5, 6])
def test_slump_frames(gpu, shape):
    w = 96
    s, st = spiky_scene(w)
    g, o = make_pair(w, w, s, st)
    g.flux_shape(shape)
    run_pair(g, o, 12, check_every=4, tol=1e-5, tag=f"slump shape {shape}")
    used = list(g.stats().flux_shape_cell_substeps)
    # the fused whole-substep shape carries the nose into its launch; with
    # the slump passes between the substeps a frame runs shape 4 instead
    if shape == 5:
        assert used[5] == 0 and used[4] == sum(used), used
    elif shape >= 0:
        assert used[shape] == sum(used), used


def test_slump_cell_index(gpu):
    w = 96
    s, st = spiky_scene(w, cell_n=True)
    g, o = make_pair(w, w, s, st)
    run_pair(g, o, 12, check_every=4, tol=1e-5, tag="slump cell n")


def test_slump_jet(gpu):
    """The jet (src/main.rs:1410-1432) with the slump on: a column builds up
    under the nozzle and sheds as it grows."""
    w, h = 128, 85
    s = scenes.jet_settings()
    s.slump = 1
    g, o = make_pair(w, h, s)
    dx = g.dx

    def deps(f, side):
        return scenes.jet_deposits(f + 1, dx, make_capsule=None if side == "g" else oracle_capsule)
    run_pair(g, o, 40, deps, check_every=10, tol=1e-5, tag="slump jet")
    rg, ro = g.mass_report(), o.mass_report()
    assert rg.deposited_ml == ro.deposited_ml
    assert abs(rg.err_pct - ro.err_pct) < 1e-6


def test_batch_refuses_slump(gpu):
    w = 64
    s, st = spiky_scene(w)
    f = czf.Film(w, w, s)
    b = czf.Batch(2)
    with pytest.raises(czf.FilmError) as e:
        b.add(f)
    assert "slump" in str(e.value)
