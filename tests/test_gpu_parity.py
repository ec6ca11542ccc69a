"""Parity of the HIP film step (through the C ABI) with the CPU restatement on
the same seeded inputs.

Bar: height and pigments within 1e-5 relative after one frame and 1e-3
after long runs (both with a floor of DRAINED_FILM = 1e-9 m for heights);
tile flags, active-tile lists, nose offsets (integer bits), lay tags and the
cure masks bit-exact; the pool8 maxima within the height tolerance.

The blob scenes run at the reference's Newtonian index n = 1 and at the
shear-thinning n = 0.7 and 0.4, where hb_flux and band_flux take their
power-law branches (film.wgsl:94-126), and with an index per cell drawn in
[0.4, 1] (settings.n_per_cell), where the cell indices are compared too.
"""
import numpy as np
import pytest

import scenes
from helpers import make_pair, oracle_capsule, rel_err

pytestmark = pytest.mark.gpu

H_FLOOR = 1e-9
N_INDEX = [1.0, 0.7, 0.4]


def compare(g, o, tol, tag=""):
    hg, ho = g.buffer("h"), o.buf("h")
    e_h = rel_err(hg, ho, H_FLOOR)
    assert e_h <= tol, f"{tag}: h rel err {e_h:.3e} (bit-equal {np.mean(hg == ho):.6f})"
    sg, so = g.buffer("strata_h"), o.buf("strata_h")
    e_s = rel_err(sg, so, H_FLOOR)
    assert e_s <= tol, f"{tag}: strata_h rel err {e_s:.3e}"
    # pigments: f16 storage, compare decoded values
    pg = g.buffer("strata_pig").view(np.float16).astype(np.float32)
    po = o.buf("strata_pig").view(np.float16).astype(np.float32)
    e_p = rel_err(pg, po, 1e-3)
    assert e_p <= max(tol, 1e-3 if tol >= 1e-3 else tol), f"{tag}: pigment rel err {e_p:.3e}"
    rg, ro = g.buffer("rheo"), o.buf("rheo")
    e_r = rel_err(rg, ro, 1e-6)
    assert e_r <= max(tol, 1e-5) * 10, f"{tag}: rheo rel err {e_r:.3e}"
    if o.settings.n_per_cell:
        e_n = rel_err(g.buffer("n_cell"), o.buf("n_cell"), 1e-6)
        assert e_n <= max(tol, 1e-5), f"{tag}: cell index rel err {e_n:.3e}"
    return e_h


def compare_masks(g, o, tag=""):
    tf_g, tf_o = g.buffer("tile_flags"), o.buf("tile_flags")
    assert np.array_equal(tf_g, tf_o), f"{tag}: tile flags differ at {np.flatnonzero(tf_g != tf_o)[:8]}"
    n = o.active_count()
    assert np.array_equal(g.buffer("tile_list")[:n], o.buf("tile_list")[:n]), f"{tag}: tile list differs"
    assert g.stats().active_tiles == n
    dg, do = g.buffer("delta"), o.buf("delta")
    assert np.array_equal(dg, do), f"{tag}: nose offsets differ at {np.argwhere(dg != do)[:8].tolist()}"
    assert np.array_equal(g.buffer("lay"), o.buf("lay")), f"{tag}: lay differs"
    assert np.array_equal(g.buffer("tile_moved"), o.buf("tile_moved")), f"{tag}: tile_moved differs"
    # cure mask: which strata hit cure >= 1 (the commit test of cure.wgsl:205)
    cg, co = g.buffer("strata_cure"), o.buf("strata_cure")
    assert np.array_equal(cg >= 1.0, co >= 1.0), f"{tag}: cure mask differs"
    # pool8 (film.wgsl:1304-1324): 4x4 maxima of the surface and the relief
    pg, po = g.pool(), o.buf("pool")[: g.pool().size]
    e_pool = rel_err(pg, po, H_FLOOR)
    assert e_pool <= 1e-5, f"{tag}: pool8 rel err {e_pool:.3e}"


def run_pair(g, o, frames, deposits_of=None, check_every=0, tol=1e-5, tag=""):
    for f in range(frames):
        dg = deposits_of(f, "g") if deposits_of else []
        do = deposits_of(f, "o") if deposits_of else []
        g.step(dg)
        o.step(do)
        if check_every and (f + 1) % check_every == 0:
            compare(g, o, tol, f"{tag} frame {f + 1}")
            compare_masks(g, o, f"{tag} frame {f + 1}")
    g.sync()


def blob_settings_n(w, tilt_deg, n_index):
    s = scenes.blobs_settings(w, tilt_deg=tilt_deg)
    s.n_index = n_index
    return s


@pytest.mark.parametrize("n_index", N_INDEX)
def test_blobs_one_frame(gpu, n_index):
    w = h = 128
    s = blob_settings_n(w, 15.0, n_index)
    st = scenes.blobs_state(w, h, seed=11, nblobs=10, sigma_cells=(2.0, 10.0))
    g, o = make_pair(w, h, s, st)
    g.step()
    o.step()
    g.sync()
    compare(g, o, 1e-5, f"blobs 1 frame n {n_index}")
    compare_masks(g, o, f"blobs 1 frame n {n_index}")
    assert abs(g.params.dt - o.params.dt) == 0.0


@pytest.mark.parametrize("n_index", N_INDEX)
def test_blobs_frames(gpu, n_index):
    w = h = 128
    s = blob_settings_n(w, 15.0, n_index)
    st = scenes.blobs_state(w, h, seed=12, nblobs=12, sigma_cells=(2.0, 10.0))
    g, o = make_pair(w, h, s, st)
    run_pair(g, o, 20, check_every=5, tol=1e-5, tag=f"blobs n {n_index}")
    rg, ro = g.mass_report(), o.mass_report()
    assert abs(rg.wet_ml - ro.wet_ml) <= 1e-5 * abs(ro.wet_ml) + 1e-12


@pytest.mark.parametrize("n_index", N_INDEX)
def test_flowing_blobs(gpu, n_index):
    """Low yield stress and a steep tilt: the paint keeps running for the
    whole 30 frames, so the power-law flux and the strata split act on many
    faces every substep."""
    w = h = 128
    s = blob_settings_n(w, 45.0, n_index)
    st = scenes.blobs_state(w, h, seed=13, nblobs=14, sigma_cells=(3.0, 12.0), peak=(2e-4, 2e-3),
                            tau=(0.01, 0.3), mu=(0.5, 5.0))
    g, o = make_pair(w, h, s, st)
    run_pair(g, o, 30, check_every=10, tol=1e-5, tag=f"flowing n {n_index}")
    assert o.active_count() > 0, "the scene settled: the test does not exercise the flux"


@pytest.mark.parametrize("tilt", [0.0, 30.0])
def test_scripted_scene(gpu, tilt):
    """The reference's deterministic no-script scene (src/main.rs:386-419)."""
    w, h = 128, 85
    s = scenes.jet_settings()
    s.tilt_deg = tilt
    g, o = make_pair(w, h, s)
    dx = g.dx

    def deps(f, side):
        return scenes.scripted_deposits(f + 1, s.width_m, s.height_m, dx,
                                        make_capsule=None if side == "g" else oracle_capsule)
    run_pair(g, o, 80, deps, check_every=20, tol=1e-5, tag=f"scripted tilt {tilt}")
    rg, ro = g.mass_report(), o.mass_report()
    assert rg.deposited_ml == ro.deposited_ml
    assert abs(rg.err_pct - ro.err_pct) < 1e-6


def test_jet_stream(gpu):
    w, h = 128, 85
    s = scenes.jet_settings()
    g, o = make_pair(w, h, s)
    dx = g.dx

    def deps(f, side):
        return scenes.jet_deposits(f + 1, dx, make_capsule=None if side == "g" else oracle_capsule)
    run_pair(g, o, 40, deps, check_every=10, tol=1e-5, tag="jet")


@pytest.mark.parametrize("n_index", N_INDEX)
def test_implicit_cg(gpu, n_index):
    """Thick, fluid paint: the implicit hydrostatic CG runs (implicit_steps > 0)."""
    w = h = 96
    s = blob_settings_n(w, 5.0, n_index)
    st = scenes.blobs_state(w, h, seed=5, nblobs=6, sigma_cells=(6.0, 12.0), peak=(4e-3, 6e-3),
                            mu=(0.05, 0.1), tau=(1e-3, 2e-3))
    g, o = make_pair(w, h, s, st)
    for _ in range(3):
        g.step()
        o.step()
    g.sync()
    assert o.implicit_steps() > 0, "scene does not exercise the implicit solve"
    assert g.stats().implicit_steps == o.implicit_steps()
    compare(g, o, 1e-5, f"implicit n {n_index}")
    compare_masks(g, o, f"implicit n {n_index}")
    assert np.allclose(g.stats().cg_report[:], o.cg_s()[4:], rtol=1e-5, atol=0)


def test_cure_commit(gpu):
    """Strata at the end of their cure commit into relief and reflectance."""
    w = h = 64
    s = scenes.blobs_settings(w, tilt_deg=0.0)
    st = scenes.blobs_state(w, h, seed=9, nblobs=6, sigma_cells=(3.0, 8.0))
    wet = st["h"] > 0
    st["strata_cure"][wet, 0] = 0.99995
    g, o = make_pair(w, h, s, st)
    for _ in range(2):
        g.step()
        o.step()
    g.sync()
    compare(g, o, 1e-5, "cure")
    compare_masks(g, o, "cure")
    mg, mo = g.buffer("m"), o.buf("m")
    assert rel_err(mg, mo, 1e-9) <= 1e-5
    assert np.array_equal(g.buffer("sub_refl"), o.buf("sub_refl"))
    assert (mo > st["m"]).any(), "no stratum was committed"


@pytest.mark.parametrize("shape", [0, 1, 2, 3, 4, 5, 6])
def test_every_substep_shape(gpu, shape):
    """Each launch shape of the substep pinned (cz_film_flux_shape): 256- and
    1024-thread tiles, staged, face-parallel and fused flux, and the whole
    substep in one launch give the restatement's results."""
    w = h = 112
    s = blob_settings_n(w, 35.0, 0.7)
    st = scenes.blobs_state(w, h, seed=14, nblobs=12, sigma_cells=(3.0, 12.0), peak=(2e-4, 2e-3),
                            tau=(0.01, 0.3), mu=(0.5, 5.0))
    g, o = make_pair(w, h, s, st)
    g.flux_shape(shape)
    run_pair(g, o, 12, check_every=4, tol=1e-5, tag=f"shape {shape}")
    used = list(g.stats().flux_shape_cell_substeps)
    # the whole-substep shape runs a frame's first substep as shape 4 when
    # the frame's substep count is odd
    others = used[4] if shape == 5 else 0
    assert used[shape] > 0 and sum(used) == used[shape] + others, used


def cell_n_settings(w, tilt_deg):
    s = blob_settings_n(w, tilt_deg, 0.7)
    s.n_per_cell = 1
    return s


def test_cell_index_flowing_blobs(gpu):
    """Every wet cell with its own index in [0.4, 1]: faces between cells of
    different indices, and the indices carried by the strata as the paint
    runs, for 30 frames."""
    w = h = 128
    s = cell_n_settings(w, 45.0)
    st = scenes.blobs_state(w, h, seed=13, nblobs=14, sigma_cells=(3.0, 12.0), peak=(2e-4, 2e-3),
                            tau=(0.01, 0.3), mu=(0.5, 5.0), n_range=(0.4, 1.0))
    g, o = make_pair(w, h, s, st)
    run_pair(g, o, 30, check_every=10, tol=1e-5, tag="flowing cell n")
    assert o.active_count() > 0
    n = o.buf("n_cell")[o.buf("h") > 0]
    assert n.min() < 0.6 and n.max() > 0.8, "the scene does not mix indices"


@pytest.mark.parametrize("shape", [0, 1, 2, 3, 4, 5, 6])
def test_cell_index_every_substep_shape(gpu, shape):
    w = h = 112
    s = cell_n_settings(w, 35.0)
    st = scenes.blobs_state(w, h, seed=14, nblobs=12, sigma_cells=(3.0, 12.0), peak=(2e-4, 2e-3),
                            tau=(0.01, 0.3), mu=(0.5, 5.0), n_range=(0.4, 1.0))
    g, o = make_pair(w, h, s, st)
    g.flux_shape(shape)
    run_pair(g, o, 12, check_every=4, tol=1e-5, tag=f"cell n shape {shape}")
    used = list(g.stats().flux_shape_cell_substeps)
    others = used[4] if shape == 5 else 0
    assert used[shape] > 0 and sum(used) == used[shape] + others, used


def test_cell_index_implicit_cg(gpu):
    """The implicit hydrostatic solve's face mobilities with per-cell indices."""
    w = h = 96
    s = cell_n_settings(w, 5.0)
    st = scenes.blobs_state(w, h, seed=5, nblobs=6, sigma_cells=(6.0, 12.0), peak=(4e-3, 6e-3),
                            mu=(0.05, 0.1), tau=(1e-3, 2e-3), n_range=(0.4, 1.0))
    g, o = make_pair(w, h, s, st)
    for _ in range(3):
        g.step()
        o.step()
    g.sync()
    assert o.implicit_steps() > 0, "scene does not exercise the implicit solve"
    compare(g, o, 1e-5, "implicit cell n")
    compare_masks(g, o, "implicit cell n")
