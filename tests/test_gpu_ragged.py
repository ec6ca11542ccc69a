"""Ragged and degenerate canvas sizes against the restatement: one cell, one
row, one column, partial tiles on both axes, with blobs that reach the
edges, a 30 deg tilt and the shear-thinning index, through the tuned substep
shapes and a pinned wide one.  Masks bit-exact, heights within 1e-5."""
import pytest

import scenes
from helpers import make_pair
from test_gpu_parity import compare, compare_masks

pytestmark = pytest.mark.gpu

SIZES = [(1, 1), (1, 29), (29, 1), (17, 17), (33, 2), (2, 47), (31, 18)]


def ragged_pair(w, h, seed):
    dx = 1.2 / 2048
    s = scenes.blobs_settings(max(w, h), tilt_deg=30.0, dx=dx)
    s.width_m = w * dx
    s.height_m = h * dx
    s.n_index = 0.7
    st = scenes.blobs_state(w, h, seed=seed, nblobs=3, sigma_cells=(1.0, 4.0), peak=(2e-4, 2e-3),
                            tau=(0.01, 0.3), mu=(0.5, 5.0))
    return make_pair(w, h, s, st)


@pytest.mark.parametrize("shape", [-1, 6])
@pytest.mark.parametrize("w,h", SIZES)
def test_ragged_canvas(gpu, w, h, shape):
    g, o = ragged_pair(w, h, seed=w * 100 + h)
    g.flux_shape(shape)
    for f in range(8):
        g.step()
        o.step()
        if f % 4 == 3:
            tag = f"{w}x{h} shape {shape} frame {f + 1}"
            compare(g, o, 1e-5, tag)
            compare_masks(g, o, tag)
    g.sync()
