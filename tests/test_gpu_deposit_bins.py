"""upload_deposits' tile binning on the device (k_dep_rects, k_dep_scan,
k_dep_fill) gives the host's bins (src/film/mod.rs:1175-1224): the same
per-tile starts, counts and ascending items, also past the item cap, and the
same frames."""
import numpy as np
import pytest

import canvaz_film as czf
import scenes

pytestmark = pytest.mark.gpu

PARTS = ("h", "strata_h", "strata_pig", "strata_cure", "rheo", "lay", "tile_flags")


def shifted(d, p0, p1):
    e = czf.Deposit.from_buffer_copy(d)
    e.p0[0], e.p0[1] = p0
    e.p1[0], e.p1[1] = p1
    return e


def deposits(rng, n, w_m, h_m, dx, r_cells=(1.0, 40.0)):
    """n capsules at random places (some entirely or partly outside the
    canvas), cloned from a few templates of radius in r_cells: the binning
    reads only positions and radii."""
    pig = scenes.jet_paint()
    radii = np.exp(np.linspace(np.log(r_cells[0]), np.log(r_cells[1]), 6)) * dx
    templates = [czf.deposit_capsule((0.0, 0.0), (0.0, 0.0), float(r), 1e-9, pig, dx) for r in radii]
    out = []
    for _ in range(n):
        p0 = (rng.uniform(-0.2, 1.2) * w_m, rng.uniform(-0.2, 1.2) * h_m)
        ang = rng.uniform(0, 2 * np.pi)
        ln = rng.uniform(0, 30) * dx
        p1 = (p0[0] + ln * np.cos(ang), p0[1] + ln * np.sin(ang))
        out.append(shifted(templates[rng.integers(len(templates))], p0, p1))
    return out


def bins(f):
    cnt = f.buffer("dep_count")
    start = f.buffer("dep_start")
    total = int(cnt.sum())
    return start, cnt, f.buffer("dep_items")[:total]


def test_device_bins_equal_host_bins(gpu):
    w, h = 320, 256
    s = scenes.blobs_settings(w, tilt_deg=20.0)
    s.height_m = h * s.width_m / w
    st = scenes.blobs_state(w, h, seed=21, nblobs=10, sigma_cells=(2.0, 12.0))
    host, dev = czf.Film(w, h, s), czf.Film(w, h, s)
    host.bin_on_device(0)
    dev.bin_on_device(1)
    for f in (host, dev):
        assert f.write_state_arrays(st)
    rng = np.random.default_rng(5)
    w_m, h_m, dx = s.width_m, s.height_m, host.dx
    frames = [deposits(rng, 3000, w_m, h_m, dx), deposits(rng, 5, w_m, h_m, dx), [],
              deposits(rng, 2000, w_m, h_m, dx, r_cells=(50.0, 200.0))]
    for k, deps in enumerate(frames):
        host.step(deps)
        dev.step(deps)
        host.sync()
        dev.sync()
        if deps:
            for a, b, name in zip(bins(host), bins(dev), ("start", "count", "items")):
                assert np.array_equal(a, b), f"frame {k}: {name}"
        for name in PARTS:
            assert np.array_equal(host.buffer(name).view(np.uint32), dev.buffer(name).view(np.uint32)), \
                f"frame {k}: {name}"


def test_device_bins_past_the_item_cap(gpu):
    """8192 thin diagonal capsules whose rectangles cover nearly every
    64-tile: more (deposit, tile) pairs than CZ_MAX_TILE_ITEMS (65536 x 16),
    so the last tiles' counts are cut as on the host."""
    w = h = 1024
    s = scenes.blobs_settings(w, tilt_deg=0.0)
    host, dev = czf.Film(w, h, s), czf.Film(w, h, s)
    host.bin_on_device(0)
    dev.bin_on_device(1)
    dx = host.dx
    rng = np.random.default_rng(9)
    tpl = czf.deposit_capsule((0.0, 0.0), (0.0, 0.0), 0.5 * dx, 1e-15, scenes.jet_paint(), dx)
    span = 0.95 * w * dx
    deps = []
    for _ in range(8192):
        # wholly inside the canvas: the host's volume accounting then takes no
        # lattice samples (inside_fraction of a long capsule crossing an edge
        # samples its whole bounding box)
        x, y = rng.uniform(2 * dx, 0.04 * w * dx), rng.uniform(2 * dx, 0.04 * h * dx)
        deps.append(shifted(tpl, (x, y), (x + span, y + span)))
    host.upload_deposits(deps)
    dev.upload_deposits(deps)
    for f in (host, dev):
        f.encode_frame(4)
        f.sync()
    a, b = bins(host), bins(dev)
    assert int(a[1].sum()) == 65536 * 16  # CZ_MAX_TILE_ITEMS: the cap was reached
    for x, y, name in zip(a, b, ("start", "count", "items")):
        assert np.array_equal(x, y), name
