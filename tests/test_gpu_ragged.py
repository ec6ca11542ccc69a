"""Ragged and degenerate canvas sizes against the restatement: one cell, one
row, one column, partial tiles on both axes, with blobs that reach the
edges, a 30 deg tilt and the shear-thinning index, through the tuned substep
shapes and a pinned wide one (masks bit-exact, heights within 1e-5), and the
same canvases as members of one batch (bit-identical to each stepped alone)."""
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


def test_ragged_members_of_a_batch(gpu):
    """The same ragged canvases as members of one batch: each ends every
    checked frame bit-identical to its twin stepped alone."""
    import canvaz_film as czf
    from test_gpu_batch import assert_same

    def film(w, h):
        dx = 1.2 / 2048
        s = scenes.blobs_settings(max(w, h), tilt_deg=30.0, dx=dx)
        s.width_m = w * dx
        s.height_m = h * dx
        s.n_index = 0.7
        st = scenes.blobs_state(w, h, seed=w * 100 + h, nblobs=3, sigma_cells=(1.0, 4.0), peak=(2e-4, 2e-3),
                                tau=(0.01, 0.3), mu=(0.5, 5.0))
        f = czf.Film(w, h, s)
        assert f.write_state_arrays(st)
        return f

    alone = [film(w, h) for w, h in SIZES]
    inb = [film(w, h) for w, h in SIZES]
    batch = czf.Batch(len(SIZES))
    for f in inb:
        batch.add(f)
    for frame in range(8):
        for f in alone:
            f.step()
        batch.step()
        if frame % 4 == 3:
            batch.sync()
            for k, (w, h) in enumerate(SIZES):
                alone[k].sync()
                assert_same(alone[k], inb[k], f"{w}x{h} frame {frame + 1}")
