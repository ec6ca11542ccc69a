"""Deposit geometry, colour tables and host-side parameters.

- The reference's own deposit tests (src/film/mod.rs:1520-1571): 300 strokes
  from its LCG whose lattice volume must equal, bit for bit, the sum of an
  independent restatement of the height profile (here in numpy float32), and
  six ribbons/capsules whose fine-grid volume is within 3 % of `vol`.
- The deposit kernel's 4x4 sub-samples coincide with the host lattice, so the
  volume a frame adds equals `vol` for an interior deposit (README.md:75).
- default_pigments / pigment_table against the reference's assets/pigments.ron
  (tests/golden/pigments_ron.json); cmf_table against the reference's CMF
  and D65 data (tests/golden/cmf_d65.json).
- The product's host functions (libcanvaz_film.so) against the restatement:
  update_params, choose_substeps, deposits, weave, substrate, bit for bit.
"""
import ctypes as C
import json
import math
import os

import numpy as np
import pytest

import canvaz_film as czf
import oracle as orc
import scenes
from cz_types import Deposit, default_overrides, default_settings
from helpers import make_pair, oracle_capsule

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
OL = orc.lib()
f32 = np.float32


def lcg_stream(seed=12345):
    """rnd() of src/film/mod.rs:1524-1528."""
    state = seed

    def rnd():
        nonlocal state
        state = (state * 6364136223846793005 + 1442695040888963407) & 0xFFFFFFFFFFFFFFFF
        return f32(state >> 40) / f32(1 << 24)
    return rnd


def reference_height_np(d, px, py):
    """reference_height (src/film/mod.rs:1487-1518) in numpy float32 arithmetic."""
    p0 = (f32(d.p0[0]), f32(d.p0[1]))
    p1 = (f32(d.p1[0]), f32(d.p1[1]))
    dx, dy = p1[0] - p0[0], p1[1] - p0[1]
    ln = np.sqrt(dx * dx + dy * dy)
    if ln > 0:
        dirx, diry, l = dx / ln, dy / ln, ln
    else:
        dirx, diry, l = f32(1), f32(0), f32(0)
    q0 = (px - p0[0], py - p0[1])
    q1 = (px - p1[0], py - p1[1])
    r0, r1, h0 = f32(d.r0), f32(d.r1), f32(d.h0)
    if l <= 0:
        x = np.sqrt(q0[0] * q0[0] + q0[1] * q0[1]) / r0
        return h0 * (f32(1) - x * x) if x < 1 else f32(0)
    cs = (d.kind & 2) != 0
    ce = (d.kind & 4) != 0
    if not cs and q0[0] * f32(d.n0[0]) + q0[1] * f32(d.n0[1]) < 0:
        return f32(0)
    if not ce and q1[0] * f32(d.n1[0]) + q1[1] * f32(d.n1[1]) >= 0:
        return f32(0)
    u = q0[0] * dirx + q0[1] * diry
    if u < -r0 or u > l + r1:
        return f32(0)
    dist = abs(q0[0] * diry - q0[1] * dirx)
    if cs and u < 0:
        dist = np.sqrt(q0[0] * q0[0] + q0[1] * q0[1])
    if ce and u > l:
        dist = np.sqrt(q1[0] * q1[0] + q1[1] * q1[1])
    t = min(max(u / l, f32(0)), f32(1))
    rt = r0 + (r1 - r0) * t
    x = dist / rt
    return h0 * (rt / r0) * (f32(1) - x * x) if x < 1 else f32(0)


def test_lattice_volume_matches_reference():
    """lattice_volume_matches_reference (src/film/mod.rs:1520-1552), 300 cases."""
    pig = np.zeros(8, np.uint32)
    dxc = f32(3.9e-4)
    rnd = lcg_stream()
    with np.errstate(all="ignore"):
        for case in range(300):
            p0 = (f32(0.1) + f32(0.05) * rnd(), f32(0.1) + f32(0.05) * rnd())
            p1 = (p0[0] + f32(4e-3) * (rnd() - f32(0.5)), p0[1] + f32(4e-3) * (rnd() - f32(0.5)))
            r = f32(2e-4) + f32(2e-3) * rnd()
            d = czf.deposit_stroke(p0, p1, float(r), 1e-9, pig, float(dxc))
            do = Deposit()
            OL.of_deposit_stroke((C.c_float * 2)(*p0), (C.c_float * 2)(*p1), float(r), 1e-9,
                                 (C.c_uint32 * 8)(*pig), float(dxc), C.byref(do))
            assert bytes(d) == bytes(do), f"case {case}: product and restatement build different deposits"
            lat = f32(0.25) * dxc
            rm = f32(max(d.r0, d.r1)) + lat
            x0 = int(np.floor((f32(min(d.p0[0], d.p1[0])) - rm) / lat))
            x1 = int(np.ceil((f32(max(d.p0[0], d.p1[0])) + rm) / lat))
            y0 = int(np.floor((f32(min(d.p0[1], d.p1[1])) - rm) / lat))
            y1 = int(np.ceil((f32(max(d.p0[1], d.p1[1])) + rm) / lat))
            s = f32(0)
            for i in range(x0, x1):
                for j in range(y0, y1):
                    s = s + reference_height_np(d, (f32(i) + f32(0.5)) * lat, (f32(j) + f32(0.5)) * lat)
            ref = s * lat * lat
            got_p = czf.lib().cz_deposit_sampled_volume(C.byref(d), float(lat))
            got_o = OL.of_sampled_volume(C.byref(do), float(lat))
            assert f32(got_p).view(np.uint32) == ref.view(np.uint32), case
            assert f32(got_o).view(np.uint32) == ref.view(np.uint32), case
            for _ in range(50):
                p = (p0[0] + f32(6e-3) * (rnd() - f32(0.5)), p0[1] + f32(6e-3) * (rnd() - f32(0.5)))
                hr = reference_height_np(d, p[0], p[1])
                hp = f32(czf.lib().cz_deposit_height_at(C.byref(d), float(p[0]), float(p[1])))
                assert hp.view(np.uint32) == f32(hr).view(np.uint32)


def fine_volume(d):
    """sampled_volume of the reference test (src/film/mod.rs:1469-1485): the
    same float32 walk (x += step) over [lo, hi) and float32 accumulation."""
    r = f32(max(d.r0, d.r1))
    step = r / f32(40)
    lo = (f32(min(d.p0[0], d.p1[0])) - f32(2) * r, f32(min(d.p0[1], d.p1[1])) - f32(2) * r)
    hi = (f32(max(d.p0[0], d.p1[0])) + f32(2) * r, f32(max(d.p0[1], d.p1[1])) + f32(2) * r)
    tot = f32(0)
    x = lo[0]
    while x < hi[0]:
        y = lo[1]
        while y < hi[1]:
            tot = tot + f32(czf.lib().cz_deposit_height_at(C.byref(d), float(x), float(y)))
            y = y + step
        x = x + step
    return float(tot * step * step)


def test_ribbon_segments_hold_their_volume():
    """ribbon_segments_hold_their_volume (src/film/mod.rs:1554-1571)."""
    pig = np.zeros(8, np.uint32)
    c, s = math.cos(0.3), math.sin(0.3)
    cases = [
        czf.deposit_ribbon((0, 0), (1e-3, 0), (1, 0), (1, 0), 4e-4, 6e-4, 0, 1e-9, pig, 3.9e-4),
        czf.deposit_ribbon((0, 0), (1e-3, 0), (c, -s), (c, s), 4e-4, 6e-4, 0, 1e-9, pig, 3.9e-4),
        czf.deposit_ribbon((0, 0), (3e-4, 0), (c, s), (c, -s), 8e-4, 8e-4, 0, 1e-9, pig, 3.9e-4),
        czf.deposit_capsule((0, 0), (2e-3, 1e-3), 5e-4, 1e-9, pig, 3.9e-4),
        czf.deposit_capsule((0, 0), (0, 0), 5e-4, 1e-9, pig, 3.9e-4),
        czf.deposit_stroke((0, 0), (4e-3, 0), 5e-4, 1e-9, pig, 3.9e-4),
    ]
    for i, d in enumerate(cases):
        v = fine_volume(d)
        assert abs(v - d.vol) < 0.03 * d.vol, f"case {i}: {v} vs {d.vol}"


def test_deposit_adds_its_volume():
    """The kernel's 4x4 sub-samples sit on the host lattice: an interior deposit adds vol."""
    w, h = 96, 64
    s = scenes.jet_settings()
    _, o = make_pair(w, h, s, gpu=False)
    dx = o.params.dx
    pig = scenes.jet_paint()
    for p0, p1, r in [((0.4, 0.3), (0.45, 0.33), 3e-3), ((0.6, 0.4), (0.6, 0.4), 5e-3), ((0.3, 0.5), (0.5, 0.52), 2e-3)]:
        d = oracle_capsule(p0, p1, r, 1e-6, pig, dx)
        before = o.buf("h").astype(np.float64).sum()
        o.upload_deposits([d])
        o.encode_frame(4, 0, 0, 0)
        added = (o.buf("h").astype(np.float64).sum() - before) * float(dx) * float(dx)
        assert added == pytest.approx(1e-6, rel=2e-4)


def test_default_pigments_match_reference_asset():
    ref = json.load(open(os.path.join(GOLDEN, "pigments_ron.json")))["pigments"]
    refl = np.zeros(12 * 16, np.float32)
    sc, ti, vi, yi = (np.zeros(12, np.float32) for _ in range(4))
    czf.lib().cz_default_pigments(czf._fp(refl), czf._fp(sc), czf._fp(ti), czf._fp(vi), czf._fp(yi))
    for i, p in enumerate(ref):
        assert np.array_equal(refl[i * 16:(i + 1) * 16], np.float32(p["reflectance"])), p["name"]
        assert (sc[i], ti[i], vi[i], yi[i]) == (p["scatter"], p["tint"], p["viscosity"], p["yield_stress"])
    tab = czf.default_pigment_table()
    r = np.clip(np.float32(ref[5]["reflectance"]), 0.01, 0.98)
    assert np.allclose(tab[5 * 40:5 * 40 + 16], (1 - r) ** 2 / (2 * r) * ref[5]["scatter"], rtol=1e-6)
    assert tab[5 * 40 + 34] == ref[5]["tint"]


def test_cmf_table():
    """cmf_table (src/color/mod.rs:145-168) from the reference's CMF and D65
    data (fixture tests/golden/cmf_d65.json): band b sums samples 2b and
    2b + 1 weighted by D65 in f32, normalised by the f32 sum of y over the
    bands.  A swapped column, a shifted sample or another illuminant fails."""
    g = json.load(open(os.path.join(GOLDEN, "cmf_d65.json")))
    cols = [np.array(g[k], np.float32) for k in ("xbar", "ybar", "zbar")]
    d65 = np.array(g["d65"], np.float32)
    want = np.zeros((16, 3), np.float32)
    ysum = f32(0.0)
    for b in range(16):
        for q in range(3):
            acc = f32(0.0)
            for s in range(2):
                acc = f32(acc + f32(cols[q][2 * b + s] * d65[2 * b + s]))
            want[b, q] = acc
        ysum = f32(ysum + want[b, 1])
    want = (want / ysum).astype(np.float32)
    t = czf.cmf_table().reshape(16, 3)
    assert np.array_equal(t, want)
    assert t[:, 1].astype(np.float64).sum() == pytest.approx(1.0, rel=1e-6)
    o = np.zeros(48, np.float32)
    OL.of_cmf_table(o.ctypes.data_as(C.POINTER(C.c_float)))
    assert np.array_equal(o.reshape(16, 3), want)


def test_host_params_match_restatement():
    rng = np.random.default_rng(21)
    for trial in range(40):
        s = default_settings()
        s.tilt_deg = float(rng.uniform(0, 60))
        s.sigma = float(rng.uniform(0.02, 0.05))
        s.capillary_implicit = int(rng.random() < 0.5)
        s.width_m = float(rng.uniform(0.05, 1.5))
        w, h = int(rng.integers(32, 3000)), int(rng.integers(32, 2000))
        s.height_m = s.width_m * h / w
        ov = default_overrides()
        ov.no_hold = int(rng.random() < 0.2)
        hmax, mumin, pend = float(rng.uniform(0, 5e-3)), float(rng.uniform(0.05, 50)), float(rng.uniform(0, 1e-3))
        fdt, S = 1 / 60, int(rng.integers(1, 65))
        p, ni = czf.update_params_host(s, ov, w, h, fdt, S, hmax, mumin, pend)
        o = orc.OracleFilm(8, 8, s, czf.default_pigment_table(), czf.cmf_table())
        o.overrides.no_hold = ov.no_hold
        o.set_host_state(hmax, mumin, pend)
        # the restatement derives dx from its own (8-cell) grid: compare on an equal-dx film
        o2 = orc.OracleFilm(w, 1, s, czf.default_pigment_table(), czf.cmf_table()) if w * 1 < 4_000_000 else None
        o2.overrides.no_hold = ov.no_hold
        o2.set_host_state(hmax, mumin, pend)
        o2.update_params(fdt, S)
        po = o2.params
        for fld in ("dt", "frame_dt", "g", "phys", "ein", "cure", "misc", "front", "front2", "front3"):
            a, b = getattr(p, fld), getattr(po, fld)
            assert bytes(a) == bytes(b) if not isinstance(a, float) else a == b, (trial, fld)
        assert ni == o2.implicit_steps()
        assert czf.choose_substeps_host(s, w, fdt, hmax, mumin, pend) == o2.choose_substeps(fdt)
        o.close()
        o2.close()


def test_weave_and_substrate():
    w, h, dx = 64, 48, 3.9e-4
    a = np.zeros(w * h, np.float32)
    b = np.zeros(w * h, np.float32)
    czf.lib().cz_weave(w, h, dx, 25e-6, czf._fp(a))
    OL.of_weave(w, h, dx, 25e-6, b.ctypes.data_as(C.POINTER(C.c_float)))
    assert np.array_equal(a, b)
    # the formula of src/film/mod.rs:212-233 in float64 (sin differences only)
    k = 2 * np.pi / 1.4e-3
    j, i = np.mgrid[0:h, 0:w].astype(np.uint64)
    hsh = ((i * 73856093) ^ (j * 19349663)) & 0xFFFFFFFF
    hsh ^= hsh >> 13
    hsh = (hsh * 0x5bd1e995) & 0xFFFFFFFF
    hsh ^= hsh >> 15
    r = (hsh & 0xFFFF) / 65535.0 * 2 - 1
    ref = 25e-6 * (1 + 0.6 * np.sin(k * (i + 0.5) * dx) * np.sin(k * (j + 0.5) * dx) + 0.4 * r)
    assert np.allclose(a.reshape(h, w), ref, rtol=2e-4, atol=1e-10)
    sub = np.zeros(8 * 4, np.uint32)
    czf.lib().cz_substrate_refl(4, 1, czf._up(sub))
    bands = sub.reshape(4, 8)[0].view(np.float16).astype(np.float32)
    assert np.array_equal(bands, np.float32(0.50 + 0.16 * np.arange(16, dtype=np.float32) / np.float32(15)).astype(np.float16).astype(np.float32))
