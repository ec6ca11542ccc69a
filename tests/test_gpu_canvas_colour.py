"""A film on a coloured canvas (settings canvas_rgb) through the C ABI: the
substrate the film starts from is the restatement's, and the cure commit
composites the finished strata onto it (cure.wgsl:139-153) exactly as the
restatement does."""
import numpy as np
import pytest

import scenes
from helpers import make_pair, rel_err
from test_gpu_parity import compare, compare_masks

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("rgb", [[0.74, 0.56, 0.33], [0.30, 0.45, 0.25]])
def test_coloured_canvas_cure_commit(gpu, rgb):
    w = h = 64
    s = scenes.blobs_settings(w, tilt_deg=0.0)
    s.has_canvas_rgb = 1
    s.canvas_rgb[:] = rgb
    st = scenes.blobs_state(w, h, seed=9, nblobs=6, sigma_cells=(3.0, 8.0))
    g, o = make_pair(w, h, s, None)
    # the canvas as created: every cell carries the colour's substrate
    sub = g.buffer("sub_refl")
    assert np.array_equal(sub, o.buf("sub_refl"))
    assert np.all(sub == sub.reshape(-1, 8)[0])
    # paint near the end of its cure on it, then two frames
    st["sub_refl"][...] = sub.reshape(st["sub_refl"].shape)
    wet = st["h"] > 0
    st["strata_cure"][wet, 0] = 0.99995
    assert g.write_state_arrays(st)
    from helpers import write_oracle_state
    write_oracle_state(o, st)
    for _ in range(2):
        g.step()
        o.step()
    g.sync()
    compare(g, o, 1e-5, f"canvas {rgb}")
    compare_masks(g, o, f"canvas {rgb}")
    assert np.array_equal(g.buffer("sub_refl"), o.buf("sub_refl"))
    assert rel_err(g.buffer("m"), o.buf("m"), 1e-9) <= 1e-5
    assert not np.array_equal(o.buf("sub_refl"), sub), "no stratum was committed"
