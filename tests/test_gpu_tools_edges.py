"""Edges of the journal and the tool passes on the device: a reshaped film
starts with no journal, no contacts and no knife; contacts past
CZ_MAX_CONTACTS are dropped (the first ones kept) as upload_contacts does
(src/film/mod.rs:1045); a knife whose blade and sweep are parallel, or whose
reach lies off the canvas, changes nothing; the ring refuses a new budget
once it exists."""
import numpy as np
import pytest

import canvaz_film as czf
import cz_types
import scenes
from helpers import make_pair
import test_oracle_tools as tt

pytestmark = pytest.mark.gpu

PARTS = ("h", "m", "rheo", "strata_h", "strata_pig", "strata_cure", "sub_refl", "lay")


def state_of(f):
    return {k: f.buffer(k).copy() for k in PARTS}


def same(a, b):
    return all(np.array_equal(a[k].view(np.uint32), b[k].view(np.uint32)) for k in PARTS)


def blob_film(w=128, h=96, seed=3):
    s = scenes.blobs_settings(w, tilt_deg=20.0)
    s.height_m = h * s.width_m / w
    f = czf.Film(w, h, s)
    assert f.write_state_arrays(scenes.blobs_state(w, h, seed=seed, nblobs=8, sigma_cells=(3.0, 10.0)))
    return f, s


def test_reshape_starts_without_journal_or_tools(gpu):
    f, s = blob_film(128, 128)
    f.journal_budget(0)
    f.journal_gen(3)
    f.touch_rect(-1.0, -1.0, 1.0, 1.0)
    k = cz_types.Knife()
    k.a[:] = (0.01, 0.03)
    k.b[:] = (0.05, 0.03)
    k.dir[:] = (0.0, 1.0)
    k.sweep, k.height, k.active, k.bands = 0.004, 3e-5, 1, 4
    f.set_knife(k)
    f.step()
    f.sync()
    assert f.journal_head()[0] > 0
    s2 = scenes.blobs_settings(96, tilt_deg=10.0)
    s2.height_m = 64 * s2.width_m / 96
    f.reshape(96, 64, s2)
    assert f.journal_head() == (0, 64)
    assert not np.any(f.buffer("jmeta")) and not np.any(f.buffer("jmark"))
    assert f.write_state_arrays(scenes.blobs_state(96, 64, seed=4, nblobs=5, sigma_cells=(3.0, 8.0)))
    # the same state stepped by a fresh film: identical, so no knife, journal or contact pass ran
    g = czf.Film(96, 64, s2)
    assert g.write_state_arrays(scenes.blobs_state(96, 64, seed=4, nblobs=5, sigma_cells=(3.0, 8.0)))
    for _ in range(3):
        f.step()
        g.step()
    f.sync()
    g.sync()
    assert same(state_of(f), state_of(g))
    assert len(f.take_scraps()) == 0
    assert f.journal_head()[0] == 0


def test_contacts_past_the_maximum_are_dropped(gpu):
    w, h = tt.W, tt.H
    s = scenes.blobs_settings(w)
    s.height_m = h * s.width_m / w
    st = tt.flat_state(np.float32(2e-4))  # paint everywhere: any contact that lands scrapes
    g, o = make_pair(w, h, s, st)
    dx = g.dx
    far = cz_types.Contact()
    far.p0[:] = (5.0, 5.0)
    far.p1[:] = (5.0, 5.0)
    far.r = 0.001
    hit = cz_types.Contact()
    hit.p0[:] = (30 * dx, 30 * dx)
    hit.p1[:] = (60 * dx, 40 * dx)
    hit.r, hit.zb, hit.bristle = 8 * dx, 0.0, 1
    cs = [far] * 65536 + [hit]  # the hit is contact 65537: beyond the maximum, never applied
    before = state_of(g)
    g.upload_contacts(cs)
    o.upload_contacts(cs)
    g.step()
    o.step()
    g.sync()
    for k in PARTS:
        assert np.array_equal(g.buffer(k).view(np.uint32), o.buf(k).view(np.uint32)), k
    assert not np.any(g.take_reservoir())
    # with the hit first, it does scrape
    g2, _ = make_pair(w, h, s, st)
    g2.upload_contacts([hit])
    g2.step()
    g2.sync()
    assert g2.take_reservoir()[1, 0] > 0
    assert not same(state_of(g2), before)


@pytest.mark.parametrize("case", ["parallel", "off_canvas", "zero_sweep"])
def test_knife_that_reaches_no_cell(gpu, case):
    f, s = blob_film()
    g, _ = blob_film()
    dx = f.dx
    k = cz_types.Knife()
    k.a[:] = (20 * dx, 30 * dx)
    k.b[:] = (80 * dx, 30 * dx)
    k.dir[:] = (1.0, 0.0) if case == "parallel" else (0.0, 1.0)
    k.sweep = 0.0 if case == "zero_sweep" else 6 * dx
    if case == "off_canvas":
        k.a[:] = (5.0, 5.0)
        k.b[:] = (5.2, 5.0)
    k.height, k.active, k.bands = 1e-5, 1, 8
    k.fill[:] = [1.0] * 8
    f.set_knife(k)
    for _ in range(2):
        f.step()
        g.step()
    f.sync()
    g.sync()
    assert same(state_of(f), state_of(g))
    assert len(f.take_scraps()) == 0


def test_budget_is_fixed_once_the_ring_exists(gpu):
    f, _ = blob_film(64, 64)
    f.journal_budget(100 * 53 * 256 * 4)
    assert f.journal_head() == (0, 100)
    f.journal_gen(1)
    with pytest.raises(czf.FilmError, match="ring exists"):
        f.journal_budget(0)
    f.journal_gen(0)
    f.step()
    f.sync()
    assert f.journal_head() == (0, 100)
