"""The contact-line terms of the restated face flux (film.wgsl:134-295), one
function at a time on hand-built neighbourhoods, against closed forms of the
reference's definitions evaluated in float64, lattice counts enumerated
here, and the symmetries those definitions have:

- hold_level (film.wgsl:154-156): A (1 - exp(-(max(d, 1) - 0.5) k));
- cox_speed (236-244): sigma (th^3 - adv^3) / (132 mu), zero up to the
  advancing angle, saturated at th = pi once the excess reaches the span;
- wet_gain (182-188): 1 / (W_AXIS + sqrt2 W_DIAG) on an axis,
  1 / (sqrt2/2 (sqrt2 W_AXIS + W_DIAG)) on a diagonal, invariant under the
  eight lattice symmetries;
- front_normal (160-180): axis and diagonal fronts, equivariance under the
  lattice symmetries;
- front_shape (203-222): wet-cell and half-disc counts of the radius-3 disc;
- feeders (134-145), wet_threshold (224-229), cell_excess (246-265) and
  rim_translation (267-295) on constructed cells.
A dropped term, a swapped axis, a wrong sign or a wrong step index moves at
least one of these values."""
import ctypes as C
import math

import numpy as np
import pytest

import oracle as orc
import scenes
from helpers import make_pair

L = orc.lib()
W = H = 32
W_AXIS, W_DIAG_FLUX = 0.6666667, 0.23570226   # shaders/common.wgsl:96-97
S2 = 0.70710678


def film(tilt=20.0):
    s = scenes.blobs_settings(W, tilt_deg=tilt)
    _, o = make_pair(W, H, s, None, gpu=False)
    o.update_params(1.0 / 60.0, 8)
    return o


def c_of(i, j):
    return j * W + i


def pack(ox, oy):
    return np.uint32(((ox + 128) << 8) | (oy + 128) | 0x10000)


def hold64(P, d):
    return P.front[0] * (1.0 - math.exp(-(max(d, 1.0) - 0.5) * P.front[1]))


# the eight symmetries of the square lattice, as (matrix, transform of an array indexed [j, i])
SYMS = [
    (np.array([[1, 0], [0, 1]]), lambda a: a),
    (np.array([[-1, 0], [0, 1]]), lambda a: a[:, ::-1]),
    (np.array([[1, 0], [0, -1]]), lambda a: a[::-1, :]),
    (np.array([[-1, 0], [0, -1]]), lambda a: a[::-1, ::-1]),
    (np.array([[0, 1], [1, 0]]), lambda a: a.T),
    (np.array([[0, 1], [-1, 0]]), lambda a: a.T[::-1, :]),
    (np.array([[0, -1], [1, 0]]), lambda a: a.T[:, ::-1]),
    (np.array([[0, -1], [-1, 0]]), lambda a: a.T[::-1, ::-1]),
]


def test_hold_level():
    o = film()
    P = o.params
    for d in (0.0, 0.5, 1.0, 1.5, 2.0, 5.0, 20.0, 90.5):
        got = L.of_fn_hold_level(o._p, d)
        assert got == pytest.approx(hold64(P, d), rel=2e-6, abs=1e-12), d
    # below one cell the level is the one-cell level
    assert L.of_fn_hold_level(o._p, 0.0) == L.of_fn_hold_level(o._p, 1.0)


def test_cox_speed():
    o = film()
    P = o.params
    theta = math.atan(P.ein[0])
    span = max(P.front[0] - hold64(P, 1.0), 1e-6)
    adv = min(theta + P.front2[0], 3.1)

    def cox64(e, mu):
        if e <= 0:
            return 0.0
        th = theta + (math.pi - theta) * min(max(e / span, 0.0), 1.0)
        if th <= adv:
            return 0.0
        return P.phys[1] * (th ** 3 - adv ** 3) / (132.0 * max(mu, 1e-4))

    # the excess at which th reaches the advancing angle
    e_adv = (adv - theta) / (math.pi - theta) * span
    assert 0 < e_adv < span
    for e in (-1e-4, 0.0, 0.5 * e_adv, 0.999 * e_adv, 1.01 * e_adv, 0.5 * span + 0.5 * e_adv, span, 3 * span):
        for mu in (1e-6, 1e-4, 0.3, 7.0):
            got = L.of_fn_cox_speed(o._p, e, mu)
            want = cox64(e, mu)
            assert got == pytest.approx(want, rel=3e-5, abs=1e-30), (e, mu, got, want)
    # zero up to the advancing angle; saturated beyond the span; 1/mu
    assert L.of_fn_cox_speed(o._p, 0.999 * e_adv, 1.0) == 0.0
    assert L.of_fn_cox_speed(o._p, 5 * span, 2.0) == L.of_fn_cox_speed(o._p, span, 2.0)
    assert L.of_fn_cox_speed(o._p, span, 2.0) == pytest.approx(L.of_fn_cox_speed(o._p, span, 1.0) / 2, rel=1e-6)
    sat = P.phys[1] * (math.pi ** 3 - adv ** 3) / 132.0
    assert L.of_fn_cox_speed(o._p, span, 1.0) == pytest.approx(sat, rel=3e-5)


def test_wet_gain():
    assert L.of_fn_wet_gain(1.0, 0.0) == pytest.approx(1.0 / (W_AXIS + math.sqrt(2) * W_DIAG_FLUX), rel=1e-6)
    want_diag = 1.0 / (S2 * (math.sqrt(2) * W_AXIS + W_DIAG_FLUX))
    assert L.of_fn_wet_gain(S2, S2) == pytest.approx(want_diag, rel=1e-6)
    rng = np.random.default_rng(1)
    for a in rng.uniform(0, 2 * math.pi, 50):
        n = np.array([math.cos(a), math.sin(a)])
        g0 = L.of_fn_wet_gain(float(n[0]), float(n[1]))
        for M, _ in SYMS:
            m = M @ n
            assert L.of_fn_wet_gain(float(m[0]), float(m[1])) == pytest.approx(g0, rel=2e-6)
    # the zero vector is guarded (max(., 1e-3) twice)
    assert L.of_fn_wet_gain(0.0, 0.0) == pytest.approx(1e6, rel=1e-5)


def test_front_normal_axis_and_equivariance():
    o = film()
    wet = 1e-4
    v = (C.c_float * 2)()
    # dry below row 10: cells at rows 10 and 11 see the dry side straight down
    h = np.full((H, W), wet, np.float32)
    h[:10, :] = 0.0
    o.buf("h")[...] = h
    for j in (10, 11):
        L.of_fn_front_normal(o._p, c_of(16, j), v)
        assert v[0] == pytest.approx(0.0, abs=1e-6) and v[1] == pytest.approx(-1.0, abs=1e-6), (j, v[0], v[1])
    L.of_fn_front_normal(o._p, c_of(16, 12), v)   # the radius-sqrt5 disc sees no dry cell
    assert (v[0], v[1]) == (0.0, 0.0)
    # a dry quadrant: the normal of its corner neighbour is the diagonal
    h = np.full((H, W), wet, np.float32)
    h[:10, :10] = 0.0
    o.buf("h")[...] = h
    L.of_fn_front_normal(o._p, c_of(10, 10), v)
    assert v[0] == pytest.approx(-S2, abs=1e-6) and v[1] == pytest.approx(-S2, abs=1e-6)
    # random dry patterns: the normal turns with the pattern (the cell's image
    # under each symmetry about the grid centre 15.5, 15.5)
    rng = np.random.default_rng(2)
    for trial in range(20):
        base = np.where(rng.random((H, W)) < 0.35, 0.0, wet).astype(np.float32)
        base[16, 15] = wet
        n0 = None
        for M, tr in SYMS:
            o.buf("h")[...] = np.ascontiguousarray(tr(base))
            p = M @ np.array([15 - 15.5, 16 - 15.5])
            i, j = int(round(p[0] + 15.5)), int(round(p[1] + 15.5))
            assert o.buf("h")[j, i] == wet
            L.of_fn_front_normal(o._p, c_of(i, j), v)
            got = np.array([v[0], v[1]])
            if n0 is None:
                n0 = got
            else:
                assert np.allclose(got, M @ n0, atol=2e-6), (trial, M.tolist(), got, M @ n0)


def test_front_shape_counts():
    o = film()
    P = o.params
    wet = 1e-4
    disc = [(di, dj) for dj in range(-3, 4) for di in range(-3, 4) if (di, dj) != (0, 0) and di * di + dj * dj <= 9]
    assert len(disc) == 28
    rng = np.random.default_rng(3)
    for trial in range(30):
        h = np.where(rng.random((H, W)) < 0.4, 0.0, wet).astype(np.float32)
        o.buf("h")[...] = h
        i, j = int(rng.integers(0, W)), int(rng.integers(0, H))
        a = rng.uniform(0, 2 * math.pi)
        n = (np.float32(math.cos(a)), np.float32(math.sin(a)))
        w0 = sum(1 for (di, dj) in disc if np.float32(np.float32(di) * n[0]) + np.float32(np.float32(dj) * n[1])
                 < np.float32(-0.5))
        w = sum(1 for (di, dj) in disc if 0 <= i + di < W and 0 <= j + dj < H and h[j + dj, i + di] > P.ein[3])
        want = 1.0 if w0 < 1 else min(max(w / w0, 0.0), 2.5) ** P.front2[1]
        got = L.of_fn_front_shape(o._p, c_of(i, j), float(n[0]), float(n[1]))
        assert got == pytest.approx(want, rel=2e-6), (trial, w, w0)
    # the no-shape flag (front2.w bit 2)
    P.front2[3] = 2.0
    assert L.of_fn_front_shape(o._p, c_of(5, 5), 1.0, 0.0) == 1.0


def test_feeders_and_wet_threshold():
    o = film()
    P = o.params
    rng = np.random.default_rng(4)
    for trial in range(40):
        h = rng.choice([0.0, 1e-5, 3e-5, 1e-4, 4e-4], size=(H, W)).astype(np.float32)
        o.buf("h")[...] = h
        i, j = int(rng.integers(0, W)), int(rng.integers(0, H))
        thr = np.float32(np.float32(2.0) * h[j, i] + np.float32(P.ein[3]))
        want = sum(1 for (di, dj) in ((-1, 0), (1, 0), (0, -1), (0, 1))
                   if 0 <= i + di < W and 0 <= j + dj < H and h[j + dj, i + di] > thr)
        assert L.of_fn_feeders(o._p, c_of(i, j)) == want
    # wet_threshold: the yield cap of the cell, at least the one-cell hold level
    gxy = math.hypot(P.g[0], P.g[1])
    for hc, ty in ((1e-4, 0.0), (1e-4, 5.0), (2e-5, 40.0), (0.0, 3.0), (3e-3, 0.05)):
        o.buf("h")[2, 3] = hc
        o.buf("rheo")[2, 3] = (1.0, ty)
        need = max(ty / (P.phys[0] * max(hc, 1e-4)) - gxy, 0.0)
        want = max(hold64(P, 1.0), need / max(P.g[2], 1e-3) * P.dx)
        assert L.of_fn_wet_threshold(o._p, c_of(3, 2)) == pytest.approx(want, rel=3e-6), (hc, ty)


def _excess_setup(o, ci, cj, off, far):
    """Cell (ci, cj) with offset `off`; `far` maps cells to (h, offset)."""
    o.buf("h")[...] = 0.0
    o.buf("delta")[...] = 0
    o.buf("h")[cj, ci] = 5e-4
    o.buf("delta")[cj, ci] = pack(*off)
    for (i, j), (hv, oo) in far.items():
        o.buf("h")[j, i] = hv
        o.buf("delta")[j, i] = pack(*oo) if oo is not None else 0


def _steps(off):
    """The far cells cell_excess visits: reach 4, 8, 16 against the offset,
    float32 as the reference computes them, truncated toward zero."""
    ox, oy = off
    ol = np.float32(math.hypot(ox, oy))
    inv = np.float32(-1.0) / ol
    dx, dy = np.float32(ox) * inv, np.float32(oy) * inv
    bx, by = np.float32(0.5 * np.sign(-ox)), np.float32(0.5 * np.sign(-oy))
    out = []
    for k in range(3):
        reach = np.float32(4 << k)
        out.append((int(np.float32(dx * reach + bx)), int(np.float32(dy * reach + by))))
    return out


def test_cell_excess_far_steps():
    o = film()
    P = o.params

    def ex64(hv, off):
        d = math.hypot(*off) if off is not None else math.hypot(64, 64)
        return hv - hold64(P, d if hv > 0 else 0.0)

    for off in ((-3, 0), (0, 2), (-3, -4), (2, -1), (-1, 1)):
        ci, cj = 12, 12
        st = _steps(off)
        # along the steps: heights rising, so the far ones win, until the first dry cell
        cells = [(ci + sx, cj + sy) for (sx, sy) in st]
        for dry_at in (3, 2, 1):
            far = {}
            for k, (i, j) in enumerate(cells):
                if not (0 <= i < W and 0 <= j < H):
                    continue
                far[(i, j)] = (0.0 if k >= dry_at else 1e-3 * (k + 2), (0, 1 + k))
            _excess_setup(o, ci, cj, off, far)
            want = ex64(5e-4, off)
            for k, (i, j) in enumerate(cells):
                if not (0 <= i < W and 0 <= j < H) or k >= dry_at:
                    break
                want = max(want, ex64(1e-3 * (k + 2), (0, 1 + k)))
            got = L.of_fn_cell_excess(o._p, c_of(ci, cj))
            assert got == pytest.approx(want, rel=3e-6), (off, dry_at, got, want)
    # no offset (unset): only the cell itself
    _excess_setup(o, 12, 12, (0, 0), {})
    o.buf("delta")[12, 12] = 0
    assert L.of_fn_cell_excess(o._p, c_of(12, 12)) == pytest.approx(5e-4 - hold64(P, math.hypot(64, 64)), rel=3e-6)


def test_rim_translation():
    o = film()
    P = o.params
    ts = C.c_float()
    # a wet cell a = (10, 10) at the front (offset 1 to a dry cell below), b = (11, 10) dry
    o.buf("h")[...] = 0.0
    o.buf("delta")[...] = 0
    o.buf("fnorm")[...] = 0.0
    o.buf("rheo")[...] = (1.0, 0.0)
    o.buf("h")[10, 10] = 2e-3
    o.buf("delta")[10, 10] = pack(1, 0)
    n = np.array([0.8, 0.6], np.float32)
    o.buf("fnorm")[10, 10] = n
    o.buf("rheo")[10, 10] = (0.7, 0.0)
    ca, cb = c_of(10, 10), c_of(11, 10)
    hf = 1e-3
    e = L.of_fn_cell_excess(o._p, ca)
    u = L.of_fn_cox_speed(o._p, e, 0.7)
    assert e > 0 and u > 0
    for fd, sgn in (((1.0, 0.0), 1.0), ((S2, S2), 1.0)):
        s = float(n[0] * fd[0] + n[1] * fd[1])
        got = L.of_fn_rim_translation(o._p, ca, cb, hf, fd[0], fd[1], C.byref(ts))
        # dst dry: room 1; the face share mix(1, |s|, front2.x)
        want = u * hf * 1.0 * (1.0 + (abs(s) - 1.0) * P.front2[0])
        assert got == pytest.approx(want, rel=3e-6) and ts.value == sgn
    # the face pointing back into the paint: the source is the dry cell, no flux
    assert L.of_fn_rim_translation(o._p, ca, cb, hf, -1.0, 0.0, C.byref(ts)) == 0.0
    # a wet destination with its own excess cuts the room: 1 - e_out / e
    o.buf("h")[10, 11] = 1.2e-3
    o.buf("delta")[10, 11] = pack(2, 0)
    o.buf("fnorm")[10, 11] = n
    e_out = L.of_fn_cell_excess(o._p, cb)
    assert 0 < e_out < e
    s = float(n[0])
    got = L.of_fn_rim_translation(o._p, ca, cb, hf, 1.0, 0.0, C.byref(ts))
    want = u * hf * min(max(1.0 - e_out / e, 0.0), 1.0) * (1.0 + (abs(s) - 1.0) * P.front2[0])
    assert got == pytest.approx(want, rel=3e-6)
    # both cells far from the front (beyond 3 / front.y cells): nothing
    far = int(3.0 / P.front[1]) + 2
    o.buf("delta")[10, 10] = pack(far, 0)
    o.buf("delta")[10, 11] = pack(far, 0)
    assert L.of_fn_rim_translation(o._p, ca, cb, hf, 1.0, 0.0, C.byref(ts)) == 0.0
    # the no-translation flag (front2.w bit 1)
    o.buf("delta")[10, 10] = pack(1, 0)
    P.front2[3] = 1.0
    assert L.of_fn_rim_translation(o._p, ca, cb, hf, 1.0, 0.0, C.byref(ts)) == 0.0
