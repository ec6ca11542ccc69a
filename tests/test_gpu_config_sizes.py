"""The workloads at the sizes they are measured at, against the CPU
restatement (the HIP path through the C ABI on one side, the oracle on the
other, same seeded inputs):

- the plain 2048x2048 blob canvas (64 seeded blobs, 15 deg tilt) for two
  frames, at the Newtonian n = 1, at the shear-thinning n = 0.7 and with an
  index per wet cell drawn in [0.4, 1];
- the headless jet scene at the reference's grid, 512x341 cells
  (`--auto 400 --script jet --res 512`, src/main.rs:222-223,1410-1432), for
  400 frames, with the mass report of src/main.rs:1679-1819 at the end;
- a batch of eight canvases drawn like the batched workload (side
  log-uniform in [256, 2048], tilt U[0, 60] deg, device-seeded blobs): every
  member equal to its twin stepped alone, the smallest also against the
  oracle;
- a serving trace of canvases with sides log-uniform in [512, 2048]: every
  result equal to its canvas stepped alone, the smallest against the oracle.

Bars as in test_gpu_parity: heights, strata and pigments within 1e-5
relative, masks, offsets, lists and lay tags bit-exact; batched and served
members bit-identical to the same canvas stepped alone.
"""
import math
import os
import sys

import numpy as np
import pytest

import canvaz_film as czf
import scenes
from cz_types import PART_NAMES
from helpers import make_pair, oracle_capsule, rel_err, write_oracle_state
from test_gpu_parity import compare, compare_masks

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

SAME_PARTS = ("h", "strata_h", "strata_pig", "strata_cure", "rheo", "m", "sub_refl", "lay", "delta", "fnorm",
              "tile_flags")


def state_of(f):
    return {k: f.buffer(k).copy() for k in PART_NAMES}


def assert_same(a, b, tag):
    for name in SAME_PARTS:
        x, y = a.buffer(name), b.buffer(name)
        if not np.array_equal(x.view(np.uint32), y.view(np.uint32)):
            bad = np.flatnonzero(x.view(np.uint32).ravel() != y.view(np.uint32).ravel())
            raise AssertionError(f"{tag}: {name} differs at {bad.size} words, first {bad[:6]}")


@pytest.mark.parametrize("n_index", [1.0, 0.7, "cell"])
def test_plain_2048_two_frames(gpu, n_index):
    n = 2048
    s = scenes.blobs_settings(n, tilt_deg=15.0)
    if n_index == "cell":
        s.n_per_cell = 1
        st = scenes.blobs_state(n, n, seed=2048, n_range=(0.4, 1.0))
    else:
        s.n_index = n_index
        st = scenes.blobs_state(n, n, seed=2048)
    g, o = make_pair(n, n, s, st)
    for f in range(2):
        g.step()
        o.step()
        g.sync()
        tag = f"2048^2 n {n_index} frame {f + 1}"
        print(tag, "active tiles", o.active_count(), flush=True)
        compare(g, o, 1e-5, tag)
        compare_masks(g, o, tag)
        assert g.params.dt == o.params.dt
    assert o.active_count() > 1000
    g.close()


def test_jet_512x341_400_frames(gpu):
    """src/main.rs:222-223: `--res 512` on the 1.2 m x 0.8 m canvas is a
    512 x 341 grid."""
    w, h = 512, 341
    s = scenes.jet_settings()
    g, o = make_pair(w, h, s)
    dx = g.dx
    for f in range(400):
        g.step(scenes.jet_deposits(f + 1, dx))
        o.step(scenes.jet_deposits(f + 1, dx, make_capsule=oracle_capsule))
        if (f + 1) % 100 == 0:
            g.sync()
            tag = f"jet frame {f + 1}"
            print(tag, "active tiles", o.active_count(), flush=True)
            compare(g, o, 1e-5, tag)
            compare_masks(g, o, tag)
    rg, ro = g.mass_report(), o.mass_report()
    # the mass report (src/main.rs:1812-1818): deposited volume is host
    # bookkeeping in f64 and equal; wet volume and the error from the FP64
    # sums of the two heights
    assert rg.deposited_ml == ro.deposited_ml
    assert ro.deposited_ml > 2.5
    assert abs(rg.wet_ml - ro.wet_ml) <= 1e-6 * ro.wet_ml
    assert abs(rg.err_pct - ro.err_pct) <= 1e-6
    assert abs(rg.err_pct) < 1e-3
    assert rg.nan_cells == 0 and rg.wet_cells == ro.wet_cells
    g.close()


def test_batch_of_eight_config_canvases(gpu):
    import bench
    specs = bench.batch_canvases(8, seed=4)
    sides = [w for (w, _, _) in specs]
    assert min(sides) >= 256 and max(sides) <= 2048 and len(set(sides)) > 4
    frames = 8
    alone, inb = [], []
    for (w, tilt, sd) in specs:
        for lst in (alone, inb):
            f = czf.Film(w, w, scenes.blobs_settings(w, tilt_deg=tilt))
            f.seed_blobs(czf.scene(seed=sd))
            lst.append(f)
    k_small = int(np.argmin(sides))
    w0, tilt0, _ = specs[k_small]
    alone[k_small].sync()
    _, o = make_pair(w0, w0, scenes.blobs_settings(w0, tilt_deg=tilt0), gpu=False)
    write_oracle_state(o, state_of(alone[k_small]))
    b = czf.Batch(8)
    for f in inb:
        b.add(f)
    for fr in range(frames):
        b.step()
        for f in alone:
            f.step()
        o.step()
    b.sync()
    for k, (w, tilt, sd) in enumerate(specs):
        alone[k].sync()
        assert_same(alone[k], inb[k], f"member {k} ({w}^2, tilt {tilt:.0f})")
    compare(inb[k_small], o, 1e-5, f"smallest member {w0}^2 vs oracle")
    compare_masks(inb[k_small], o, f"smallest member {w0}^2 vs oracle")
    assert sum(f.stats().active_tiles for f in inb) > 0
    b.close()
    for f in alone + inb:
        f.close()


def test_serve_trace_sides_512_to_2048(gpu):
    rng = np.random.default_rng(21)
    reqs, specs = [], []
    t = 0.0
    for k in range(10):
        w = int(round(math.exp(rng.uniform(math.log(512.0), math.log(2048.0)))))
        steps = int(rng.integers(10, 41))
        s = scenes.blobs_settings(w, tilt_deg=15.0)
        sc = czf.scene(seed=7000 + k)
        reqs.append(czf.request(k, w, w, steps, t, s, sc))
        specs.append((w, steps, s, sc))
        t += float(rng.exponential(1.0 / 200.0))
    srv = czf.Server(hbm_budget=48 << 30, max_members=6, keep_films=True)
    res, st = srv.run(reqs)
    assert all(r.status == 0 for r in res)
    assert st.peak_members > 1
    k_small = int(np.argmin([w for (w, _, _, _) in specs]))
    for r, (w, steps, s, sc) in zip(res, specs):
        f = czf.Film(w, w, s)
        f.seed_blobs(sc)
        if r.id == k_small:
            f.sync()
            _, o = make_pair(w, w, s, gpu=False)
            write_oracle_state(o, state_of(f))
            for _ in range(steps):
                o.step()
        for _ in range(steps):
            f.step()
        f.sync()
        g = srv.film(r.id, w, w)
        assert g is not None and g.stats().frame == steps
        assert_same(g, f, f"request {r.id} ({w}^2, {steps} steps)")
        if r.id == k_small:
            compare(g, o, 1e-5, f"request {r.id} vs oracle")
            compare_masks(g, o, f"request {r.id} vs oracle")
        f.close()
    srv.close()
