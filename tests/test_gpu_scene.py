"""The device blob-canvas generator equals its host twin: heights, rheology
and pigments bit for bit; the weave relief within one ulp (it takes sin in
double precision on each side); the rest of the state as write_state leaves
it."""
import numpy as np
import pytest

import canvaz_film as czf
import scenes

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("w,h,seed,n_hi", [(200, 150, 21, 0.0), (257, 64, 22, 0.0), (192, 160, 23, 1.0)])
def test_device_scene_equals_host(gpu, w, h, seed, n_hi):
    s = scenes.blobs_settings(w, tilt_deg=10.0)
    s.n_per_cell = 1 if n_hi else 0
    f = czf.Film(w, h, s)
    sc = czf.scene(seed=seed, nblobs=24, sigma_lo=2.0, sigma_hi=12.0, n_lo=0.4 if n_hi else 0.0, n_hi=n_hi)
    f.seed_blobs(sc)
    hv, mv, rv, pv = czf.scene_cells_host(w, h, f.dx, sc)
    assert np.array_equal(f.buffer("h"), hv)
    assert np.array_equal(f.buffer("rheo"), rv)
    pig = f.buffer("strata_pig")
    assert np.array_equal(pig[:, :, 0, :], pv)
    assert not pig[:, :, 1:, :].any()
    sh = f.buffer("strata_h")
    assert np.array_equal(sh[..., 0], hv) and not sh[..., 1:].any()
    assert np.allclose(f.buffer("m"), mv, rtol=2.5e-7, atol=1e-12)
    assert not f.buffer("strata_cure").any() and not f.buffer("lay").any()
    raw = np.zeros(8, np.uint32)
    czf.lib().cz_substrate_refl(1, 1, czf._up(raw))
    assert (f.buffer("sub_refl") == raw).all()
    assert (f.buffer("tile_moved") == 1).all()
    # it steps like any written state
    f.step()
    f.sync()
    assert f.stats().active_tiles > 0
